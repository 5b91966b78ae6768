// reduce.h -- deterministic block reductions shared by the iterative-solver kernels
// (256-thread workgroups).
#pragma once

#include <hip/hip_runtime.h>

namespace mg {

// Sum of v over the block: wave reductions by a fixed xor tree, then the four waves in order.
__device__ inline double block_sum_waves(double v, double *sm4) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) sm4[threadIdx.x >> 6] = v;
    __syncthreads();
    const double t = (sm4[0] + sm4[1]) + (sm4[2] + sm4[3]);
    __syncthreads();
    return t;
}

// Sum of the previous kernel's partials part[0 .. np) in a fixed order.  Every workgroup of
// the next kernel forms the same sum itself (the launch boundary makes the partials
// visible), so a scalar such as alpha or beta needs no kernel of its own and every
// workgroup holds the identical value.
__device__ inline double sum_partials(const double *__restrict__ part, int np, double *sm4) {
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    int b = threadIdx.x;
    for (; b + 3 * 256 < np; b += 4 * 256)
#pragma unroll
        for (int q = 0; q < 4; ++q) a[q] += part[b + q * 256];
    for (; b < np; b += 256) a[0] += part[b];
    return block_sum_waves((a[0] + a[1]) + (a[2] + a[3]), sm4);
}

}  // namespace mg
