// SU(2) instantiation of the lattice kernels (FP64 and FP32 forms).
#define GROUP_NS su2k
#include "tcym_kernels.inc"

namespace tcym {
Backend* make_backend_su2(int fp32) {
    if (fp32) return new su2k::Impl<SU2<float>, float>();
    return new su2k::Impl<SU2<double>, double>();
}
}  // namespace tcym
