// Final deterministic reduction of per-block FP64 partials, shared by the
// single-lattice context and the batch so that both sum in the same order.
#pragma once
#include <hip/hip_runtime.h>

namespace tcym {

// Block-wide (256 threads) reduction of quantity q of nblk partials laid out
// as part[b*nq + q]: per-thread strided Neumaier sums (or max), then a fixed
// shuffle tree and a fixed 4-wave combine.  Thread 0 returns the result.
__device__ __forceinline__ double final_reduce_block(const double* __restrict__ part, long nblk, int nq, int q, bool mx) {
    double s = 0.0, c = 0.0;
    auto add = [&](double v) {
        if (mx) s = fmax(s, v);
        else {
            const double t = s + v;
            c += fabs(s) >= fabs(v) ? (s - t) + v : (v - t) + s;
            s = t;
        }
    };
    // a thread's partials b, b + stride, ... in that order; eight loads in
    // flight at a time instead of one (the sum waited on each load's latency)
    const long stride = blockDim.x;
    long b = threadIdx.x;
    for (; b + 7*stride < nblk; b += 8*stride) {
        double v[8];
#pragma unroll
        for (int k = 0; k < 8; k++) v[k] = part[(b + k*stride)*nq + q];
#pragma unroll
        for (int k = 0; k < 8; k++) add(v[k]);
    }
    for (; b < nblk; b += stride) add(part[b*nq + q]);
    if (!mx) s += c;
    __shared__ double sm[4];
    for (int o = 32; o > 0; o >>= 1) {
        const double w = __shfl_down(s, o, 64);
        s = mx ? fmax(s, w) : s + w;
    }
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = s;
    __syncthreads();
    double r = sm[0];
    for (int w = 1; w < 4; w++) r = mx ? fmax(r, sm[w]) : r + sm[w];
    return r;
}

}  // namespace tcym
