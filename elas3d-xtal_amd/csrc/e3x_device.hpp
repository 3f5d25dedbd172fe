// Device helpers shared by the gfx950 kernels (internal).
#pragma once

#include "e3x_internal.hpp"

namespace e3x {

// Local node a of the brick sits at (PX(a), PY(a), PZ(a)):
// 1 (0,0,0) 2 (1,0,0) 3 (1,1,0) 4 (0,1,0) 5 (0,0,1) 6 (1,0,1) 7 (1,1,1) 8 (0,1,1)
// (src/elas3dxtal_pcg.f90:639, 740-765).
__host__ __device__ constexpr int PX(int a) { return ((a + 1) >> 1) & 1; }
__host__ __device__ constexpr int PY(int a) { return (a >> 1) & 1; }
__host__ __device__ constexpr int PZ(int a) { return a >> 2; }

__device__ __forceinline__ int wrap(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// Consecutive tile ids on one XCD: the hardware hands workgroup b to XCD b % 8, so
// neighbouring tiles (which share halo lines) land in the same L2.
__device__ __forceinline__ int xcd_remap(int bid, int nb)
{
    const int xcd = bid % kXcds, k = bid / kXcds;
    const int q = nb / kXcds, rem = nb % kXcds;
    return xcd * q + (xcd < rem ? xcd : rem) + k;
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// Deterministic sum over a block of NT threads (fixed shuffle tree per wave, then the
// waves in order); the result is valid in thread 0.
template <int NV, int NT = kBlock>
__device__ __forceinline__ void block_sum(double (&v)[NV])
{
    __shared__ double sh[NV][NT / 64];
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NV; q++) {
        const double t = wave_sum(v[q]);
        if (l == 0) sh[q][w] = t;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; q++) {
            double t = sh[q][0];
            for (int ww = 1; ww < NT / 64; ww++) t += sh[q][ww];
            v[q] = t;
        }
    }
}

// Block partials, layout [q][block]: plain stores, read by a separate finalize kernel
// (no device-scope fences inside the producing kernel).
template <int NV>
__device__ __forceinline__ void write_partials(double (&v)[NV], double *partials)
{
    block_sum<NV>(v);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; q++) partials[(int64_t)q * gridDim.x + blockIdx.x] = v[q];
    }
}

// Threads of the single-block finalize kernels.
constexpr int kFinBlock = 1024;

// In a finalize kernel of kFinBlock threads: fixed-order totals of the partials, valid
// in thread 0.  Each thread keeps 8 independent loads in flight.
template <int NV>
__device__ __forceinline__ void sum_partials(const double *partials, int nparts, double (&tot)[NV])
{
#pragma unroll
    for (int q = 0; q < NV; q++) {
        const double *p = partials + (int64_t)q * nparts;
        double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int i = threadIdx.x; i < nparts; i += 8 * kFinBlock) {
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int idx = i + u * kFinBlock;
                if (idx < nparts) acc[u] += p[idx];
            }
        }
        tot[q] = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
    }
    block_sum<NV, kFinBlock>(tot);
}

}  // namespace e3x
