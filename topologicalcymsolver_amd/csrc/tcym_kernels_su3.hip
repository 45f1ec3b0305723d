// SU(3) instantiation of the lattice kernels (FP64 and FP32 forms).
#define GROUP_NS su3k
#include "tcym_kernels.inc"

namespace tcym {
Backend* make_backend_su3(int fp32) {
    if (fp32) return new su3k::Impl<SU3<float>, float>();
    return new su3k::Impl<SU3<double>, double>();
}
}  // namespace tcym
