// gfx950 matrix-free stencil y = K x of the elas3dxtal_pcg solve (energy_OpenMP
// f90:1176-1321, dembx SpMV f90:1585-1884).
//
// A block owns an 8x8x4 tile of nodes and stages the 10x10x6 node halo of x and the
// 9x9x5 voxel phases in LDS; each 64-lane wave owns a 4x4x4 sub-cube (the shape most
// often inside one grain).  A node is homogeneous when its 8 voxels share one phase,
// mixed otherwise.  Each wave loops over the distinct phases it holds, in ascending
// phase order, and for phase P applies with wave-uniform coefficients read through
// the scalar cache (SGPR operands, no per-lane coefficient loads):
//   - to its homogeneous lanes of phase P the 27-point stencil of P (243 FMA; the
//     block of -delta is the transpose of the block of +delta, so 126 coefficients),
//   - to its mixed lanes, for every octant a whose voxel has phase P, the 3x24 rows of
//     the element matrix dk[P][a] (72 FMA).
// A node's sum is therefore ordered by its own phases and octants only, independent of
// the other lanes: results are bitwise reproducible.  Neighbour values are read from
// LDS at their use, which keeps registers low enough for six waves per SIMD to hide
// the scalar-cache latency.
#include "e3x_device.hpp"

#include <cmath>

namespace e3x {
namespace {

template <typename T>
struct StencilArgs {
    int nx, ny, nz, ntx, nty;
    int64_t ns;
    const uint16_t *pix;
    const T *dk, *sten;
    const T *x;            // [NR][3][ns]
    T *y;                  // [NR][3][ns] (kApply, kEnergy, kPlain)
    const double *b;       // kEnergy, kResid
    float *y32;            // kResid output
    double *partials;
    const PcgState *st;    // kApply, kResid: skip finished right-hand sides
};

// LDS offset of neighbour nb = (dz+1)*9 + (dy+1)*3 + (dx+1) relative to the centre.
__device__ constexpr int nboff(int nb) { return (nb / 9 - 1) * (HX * HY) + ((nb / 3) % 3 - 1) * HX + (nb % 3 - 1); }

template <typename T, int NR, int MODE>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(6, 8))) void k_stencil(StencilArgs<T> A)
{
    __shared__ T hs[NR * 3 * HN];
    __shared__ int ps[VN];

    unsigned live = (1u << NR) - 1u;  // right-hand sides to compute
    if constexpr (MODE == kApply || MODE == kResid) {
        live = 0u;
#pragma unroll
        for (int r = 0; r < NR; r++)
            if (A.st[r].stop_iter == 0) live |= 1u << r;
        if (!live) return;
    }

    const int bid = xcd_remap(blockIdx.x, gridDim.x);
    const int tx = bid % A.ntx, ty = (bid / A.ntx) % A.nty, tz = bid / (A.ntx * A.nty);
    const int ox = tx * TX, oy = ty * TY, oz = tz * TZ;
    const int64_t ns = A.ns;

    for (int q = threadIdx.x; q < HN; q += kBlock) {
        const int hx = q % HX, hy = (q / HX) % HY, hz = q / (HX * HY);
        const int gx = wrap(ox + hx - 1, A.nx), gy = wrap(oy + hy - 1, A.ny), gz = wrap(oz + hz - 1, A.nz);
        const int64_t gm = gx + (int64_t)A.nx * (gy + (int64_t)A.ny * gz);
#pragma unroll
        for (int rc = 0; rc < NR * 3; rc++) hs[rc * HN + q] = A.x[(int64_t)rc * ns + gm];
    }
    for (int q = threadIdx.x; q < VN; q += kBlock) {
        const int vx = q % VX, vy = (q / VX) % VY, vz = q / (VX * VY);
        const int gx = wrap(ox + vx - 1, A.nx), gy = wrap(oy + vy - 1, A.ny), gz = wrap(oz + vz - 1, A.nz);
        ps[q] = A.pix[gx + (int64_t)A.nx * (gy + (int64_t)A.ny * gz)];
    }
    __syncthreads();

    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int lx = (w & 1) * 4 + (l & 3), ly = (w >> 1) * 4 + ((l >> 2) & 3), lz = l >> 4;
    const int gx = ox + lx, gy = oy + ly, gz = oz + lz;
    const bool active = gx < A.nx && gy < A.ny && gz < A.nz;
    const int64_t m = gx + (int64_t)A.nx * (gy + (int64_t)A.ny * gz);
    int pv[8];
#pragma unroll
    for (int a = 0; a < 8; a++) pv[a] = ps[(lz + 1 - PZ(a)) * (VX * VY) + (ly + 1 - PY(a)) * VX + (lx + 1 - PX(a))];
    bool homog = true;
#pragma unroll
    for (int a = 1; a < 8; a++) homog &= pv[a] == pv[0];
    const int cidx = (lz + 1) * (HX * HY) + (ly + 1) * HX + (lx + 1);

    double part[MODE == kEnergy ? 2 * NR : NR];
#pragma unroll
    for (int q = 0; q < (MODE == kEnergy ? 2 * NR : NR); q++) part[q] = 0.0;

#pragma unroll
    for (int r = 0; r < NR; r++) {
        if (!((live >> r) & 1u)) continue;
        const T *hr = hs + (r * 3) * HN + cidx;  // component c of neighbour nb: hr[c * HN + nboff(nb)]
        T acc[3] = {T(0), T(0), T(0)};
        unsigned done = active ? 0u : 0xFFu;  // octants (mixed) or the whole node (homogeneous) applied
        while (true) {
            // smallest phase this lane still needs; the wave takes the smallest of all
            int cand = 0x7fffffff;
            if (done != 0xFFu) {
                if (homog) {
                    cand = pv[0];
                } else {
#pragma unroll
                    for (int a = 0; a < 8; a++)
                        if (!((done >> a) & 1u)) cand = min(cand, pv[a]);
                }
            }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) cand = min(cand, __shfl_xor(cand, off, 64));
            if (cand == 0x7fffffff) break;
            const int P = __builtin_amdgcn_readfirstlane(cand);
            const bool hm = homog && done != 0xFFu && pv[0] == P;
            if (__ballot(hm)) {
                const T *S = A.sten + (int64_t)P * kStenHalf;
                T t0 = T(0), t1 = T(0), t2 = T(0);
#pragma unroll
                for (int k = 0; k < 13; k++) {
                    const int np = 14 + k, nm = 12 - k;  // neighbours +delta and -delta
#pragma unroll
                    for (int n = 0; n < 3; n++) {
                        const T hp = hr[n * HN + nboff(np)];
                        t0 += S[k * 9 + 0 * 3 + n] * hp;
                        t1 += S[k * 9 + 1 * 3 + n] * hp;
                        t2 += S[k * 9 + 2 * 3 + n] * hp;
                    }
#pragma unroll
                    for (int n = 0; n < 3; n++) {  // block of -delta: S[-d][j][n] = S[d][n][j]
                        const T hq = hr[n * HN + nboff(nm)];
                        t0 += S[k * 9 + n * 3 + 0] * hq;
                        t1 += S[k * 9 + n * 3 + 1] * hq;
                        t2 += S[k * 9 + n * 3 + 2] * hq;
                    }
                }
                const T *Cc = S + 117;  // centre: 00 11 22 01 02 12
                const T c0 = hr[0], c1 = hr[HN], c2 = hr[2 * HN];
                t0 += Cc[0] * c0; t0 += Cc[3] * c1; t0 += Cc[4] * c2;
                t1 += Cc[3] * c0; t1 += Cc[1] * c1; t1 += Cc[5] * c2;
                t2 += Cc[4] * c0; t2 += Cc[5] * c1; t2 += Cc[2] * c2;
                if (hm) {
                    acc[0] += t0; acc[1] += t1; acc[2] += t2;
                    done = 0xFFu;
                }
            }
#pragma unroll
            for (int a = 0; a < 8; a++) {
                const bool mm = !homog && !((done >> a) & 1u) && pv[a] == P;
                if (__ballot(mm)) {
                    const T *D = A.dk + (int64_t)P * 576 + a * 72;
                    T t0 = T(0), t1 = T(0), t2 = T(0);
#pragma unroll
                    for (int b = 0; b < 8; b++) {
                        const int nb = (PZ(b) - PZ(a) + 1) * 9 + (PY(b) - PY(a) + 1) * 3 + (PX(b) - PX(a) + 1);
#pragma unroll
                        for (int n = 0; n < 3; n++) {
                            t0 += D[0 * 24 + b * 3 + n] * hr[n * HN + nboff(nb)];
                            t1 += D[1 * 24 + b * 3 + n] * hr[n * HN + nboff(nb)];
                            t2 += D[2 * 24 + b * 3 + n] * hr[n * HN + nboff(nb)];
                        }
                    }
                    if (mm) {
                        acc[0] += t0; acc[1] += t1; acc[2] += t2;
                        done |= 1u << a;
                    }
                }
            }
        }

        if (active) {
            if constexpr (MODE == kApply) {
                double pr = 0.0;
#pragma unroll
                for (int j = 0; j < 3; j++) {
                    A.y[((int64_t)r * 3 + j) * ns + m] = acc[j];
                    pr += (double)hr[j * HN] * (double)acc[j];
                }
                part[r] += pr;
            } else if constexpr (MODE == kEnergy) {
                double e = 0.0, gsq = 0.0;
#pragma unroll
                for (int j = 0; j < 3; j++) {
                    const int64_t q = ((int64_t)r * 3 + j) * ns + m;
                    const double uq = (double)hr[j * HN], bq = A.b[q], kx = (double)acc[j];
                    e = e + 0.5 * uq * kx + bq * uq;  // f90:1336
                    const double gq = kx + bq;         // f90:1337
                    A.y[q] = (T)gq;
                    gsq += gq * gq;                     // f90:452
                }
                part[r] += e;
                part[NR + r] += gsq;
            } else if constexpr (MODE == kResid) {
#pragma unroll
                for (int j = 0; j < 3; j++) {
                    const int64_t q = ((int64_t)r * 3 + j) * ns + m;
                    A.y32[q] = (float)((double)acc[j] + A.b[q]);
                }
            } else {
#pragma unroll
                for (int j = 0; j < 3; j++) A.y[((int64_t)r * 3 + j) * ns + m] = acc[j];
            }
        }
    }
    if constexpr (MODE == kApply || MODE == kEnergy) write_partials(part, A.partials);
}

template <int NR>
__global__ __launch_bounds__(kFinBlock) void k_energy_finish(const double *partials, int nparts, const double *cconst,
                                                          double *results)
{
    double tot[2 * NR];
    sum_partials<2 * NR>(partials, nparts, tot);
    if (threadIdx.x == 0)
        for (int r = 0; r < NR; r++) {
            results[r] = cconst[r] + tot[r];
            results[NR + r] = tot[NR + r];
        }
}

template <typename T>
StencilArgs<T> stencil_args(const Solver &s, const T *dk, const T *sten, const T *x, T *y)
{
    StencilArgs<T> A{};
    A.nx = s.nx; A.ny = s.ny; A.nz = s.nz;
    A.ntx = (s.nx + TX - 1) / TX; A.nty = (s.ny + TY - 1) / TY;
    A.ns = s.ns; A.pix = s.pix; A.dk = dk; A.sten = sten; A.x = x; A.y = y;
    A.partials = s.partials; A.st = s.st;
    return A;
}

template <typename T, int MODE>
void dispatch_stencil(const Solver &s, const StencilArgs<T> &A)
{
    const dim3 grid((unsigned)stencil_blocks(s)), block(kBlock);
    switch (s.nrhs) {
    case 1: hipLaunchKernelGGL((k_stencil<T, 1, MODE>), grid, block, 0, s.stream, A); break;
    case 2: hipLaunchKernelGGL((k_stencil<T, 2, MODE>), grid, block, 0, s.stream, A); break;
    case 3: hipLaunchKernelGGL((k_stencil<T, 3, MODE>), grid, block, 0, s.stream, A); break;
    case 4: hipLaunchKernelGGL((k_stencil<T, 4, MODE>), grid, block, 0, s.stream, A); break;
    case 5: hipLaunchKernelGGL((k_stencil<T, 5, MODE>), grid, block, 0, s.stream, A); break;
    default: hipLaunchKernelGGL((k_stencil<T, 6, MODE>), grid, block, 0, s.stream, A); break;
    }
}

}  // namespace

int64_t stencil_blocks(const Solver &s)
{
    return (int64_t)((s.nx + TX - 1) / TX) * ((s.ny + TY - 1) / TY) * ((s.nz + TZ - 1) / TZ);
}

hipError_t launch_energy(Solver &s)
{
    StencilArgs<double> A = stencil_args<double>(s, s.dk, s.sten, s.u, s.g);
    A.b = s.b;
    dispatch_stencil<double, kEnergy>(s, A);
    const int np = (int)stencil_blocks(s);
    const dim3 one(1), fin(kFinBlock);
    const double *p = s.partials;
    switch (s.nrhs) {
    case 1: hipLaunchKernelGGL(k_energy_finish<1>, one, fin, 0, s.stream, p, np, s.cconst, s.results); break;
    case 2: hipLaunchKernelGGL(k_energy_finish<2>, one, fin, 0, s.stream, p, np, s.cconst, s.results); break;
    case 3: hipLaunchKernelGGL(k_energy_finish<3>, one, fin, 0, s.stream, p, np, s.cconst, s.results); break;
    case 4: hipLaunchKernelGGL(k_energy_finish<4>, one, fin, 0, s.stream, p, np, s.cconst, s.results); break;
    case 5: hipLaunchKernelGGL(k_energy_finish<5>, one, fin, 0, s.stream, p, np, s.cconst, s.results); break;
    default: hipLaunchKernelGGL(k_energy_finish<6>, one, fin, 0, s.stream, p, np, s.cconst, s.results); break;
    }
    return hipGetLastError();
}

hipError_t launch_apply_plain(Solver &s, const double *x, double *y)
{
    StencilArgs<double> A = stencil_args<double>(s, s.dk, s.sten, x, y);
    dispatch_stencil<double, kPlain>(s, A);
    return hipGetLastError();
}

hipError_t launch_apply_hah(Solver &s, bool mixed, int it)
{
    if (mixed) dispatch_stencil<float, kApply>(s, stencil_args<float>(s, s.dk32, s.sten32, s.h32, s.ah32));
    else dispatch_stencil<double, kApply>(s, stencil_args<double>(s, s.dk, s.sten, s.h, s.ah));
    return launch_hah_finish(s, (int)stencil_blocks(s), it);
}

hipError_t launch_apply_only(Solver &s, bool mixed)
{
    if (mixed) dispatch_stencil<float, kApply>(s, stencil_args<float>(s, s.dk32, s.sten32, s.h32, s.ah32));
    else dispatch_stencil<double, kApply>(s, stencil_args<double>(s, s.dk, s.sten, s.h, s.ah));
    return hipGetLastError();
}

hipError_t launch_resid(Solver &s)
{
    StencilArgs<double> A = stencil_args<double>(s, s.dk, s.sten, s.u, nullptr);
    A.b = s.b;
    A.y32 = s.g32;
    dispatch_stencil<double, kResid>(s, A);
    return hipGetLastError();
}

}  // namespace e3x
