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

// Phase of voxel m (0-based): the ids are uint16, or uint32 on a wide handle (more than
// kNarrowPhases phases).  `wide` is uniform over a launch.
__device__ __forceinline__ int phase_at(const void *pix, int wide, int64_t m)
{
    return wide ? (int)static_cast<const uint32_t *>(pix)[m] : (int)static_cast<const uint16_t *>(pix)[m];
}

// Homogeneous phase of node m, or -1 when its 8 voxels hold several phases (the all-ones
// id in either width).
__device__ __forceinline__ int hom_phase_at(const void *hphase, int wide, int64_t m)
{
    if (wide) return (int)static_cast<const uint32_t *>(hphase)[m];
    const int v = static_cast<const uint16_t *>(hphase)[m];
    return v == 0xFFFF ? -1 : v;
}

// Consecutive tile ids on one XCD: the hardware hands workgroup b to XCD b % 8, so
// neighbouring tiles (which share halo lines) land in the same L2.
__device__ __forceinline__ int xcd_remap(int bid, int nb)
{
    const int xcd = bid % kXcds, k = bid / kXcds;
    const int q = nb / kXcds, rem = nb % kXcds;
    return xcd * q + (xcd < rem ? xcd : rem) + k;
}

// v of lane i - SHIFT within its row of 16 lanes (row_shr), 0 where that runs off the
// row; CTRL >= 0x142 are the row broadcasts of lane 15 / 31 into the rows in ROWS.
template <int CTRL, int ROWS = 0xf>
__device__ __forceinline__ double dpp_f64(double v)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROWS, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROWS, 0xf, false);
    return __hiloint2double(hi, lo);
}

// Sum over the 64 lanes of a wave (all lanes active), valid in lane 63: a fixed DPP
// tree (row scans, then row broadcasts), so the order of the additions never varies.
__device__ __forceinline__ double wave_sum(double v)
{
    v += dpp_f64<0x111>(v);        // row_shr:1
    v += dpp_f64<0x112>(v);        // row_shr:2
    v += dpp_f64<0x114>(v);        // row_shr:4
    v += dpp_f64<0x118>(v);        // row_shr:8  -> lane 15 of each row holds the row sum
    v += dpp_f64<0x142, 0xa>(v);   // row_bcast:15 into rows 1, 3
    v += dpp_f64<0x143, 0xc>(v);   // row_bcast:31 into rows 2, 3 -> lane 63 holds the total
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
        if (l == 63) sh[q][w] = t;
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

// True in every lane of the last wave of its block to call it, false in the others,
// without a block barrier: a wave that is done leaves at once instead of waiting for the
// slowest wave of its block, and the last one finds in LDS everything the others wrote
// there before their call.  `arrived` is a block-shared counter set to 0 before a
// __syncthreads() that precedes the call; every wave calls it once, all lanes active.
template <int NT>
__device__ __forceinline__ bool last_wave(unsigned *arrived)
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    unsigned prev = 0;
    if ((threadIdx.x & 63) == 63)
        prev = __hip_atomic_fetch_add(arrived, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
    prev = __builtin_amdgcn_readlane(prev, 63);
    if (prev != NT / 64 - 1) return false;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    return true;
}

// Threads of the single-block finalize kernels.
constexpr int kFinBlock = 1024;

// In a finalize kernel of kFinBlock threads: fixed-order totals of the partials, valid
// in thread 0.  Each thread keeps U loads in flight (U = 32 for one or two sums, so that
// the 31,250 partials of a 200^3 apply come in one round of memory latency; 8 for the
// wide finalizers), adds them in a fixed tree, then the block sum.
template <int NV>
__device__ __forceinline__ void sum_partials(const double *partials, int nparts, double (&tot)[NV])
{
    constexpr int U = NV <= 2 ? 32 : 8;
#pragma unroll
    for (int q = 0; q < NV; q++) {
        const double *p = partials + (int64_t)q * nparts;
        double acc[U];
#pragma unroll
        for (int u = 0; u < U; u++) acc[u] = 0.0;
        for (int i = threadIdx.x; i < nparts; i += U * kFinBlock) {
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int idx = i + u * kFinBlock;
                if (idx < nparts) acc[u] += p[idx];
            }
        }
#pragma unroll
        for (int w = U / 2; w >= 1; w /= 2)
#pragma unroll
            for (int u = 0; u < w; u++) acc[u] += acc[u + w];
        tot[q] = acc[0];
    }
    block_sum<NV, kFinBlock>(tot);
}

}  // namespace e3x
