// gfx950 matrix-free stencil y = K x of the elas3dxtal_pcg solve (energy_OpenMP
// f90:1176-1321, dembx SpMV f90:1585-1884).
//
// A block owns an 8x8x4 tile of nodes and stages the 10x10x6 node halo of x and the
// 9x9x5 voxel phases in LDS.  A node is homogeneous when its 8 voxels share one phase,
// mixed otherwise.  After staging, the block packs its homogeneous nodes into the
// leading lanes and its mixed nodes into the trailing ones (a stable ballot
// compaction), so that most waves hold one kind only.  Coefficients are wave-uniform
// and read through the scalar cache (SGPR operands, no per-lane coefficient loads):
//   - homogeneous lanes: the wave loops over the phases its lanes hold; for phase P the
//     27-point stencil of P (243 FMA; the block of -delta is the transpose of the block
//     of +delta, so 126 coefficients) goes to the lanes of phase P;
//   - mixed lanes: for each octant a in order 0..7 the wave reads the 24 neighbour
//     values of voxel a once, then loops over the phases its lanes have at octant a and
//     applies the 3x24 rows of the element matrix dk[P][a] (72 FMA) to those lanes.
// The next phase is the one of the first lane still pending (ballot + readlane).  Each
// lane is served exactly once per octant (or once in all, when homogeneous), so a
// node's sum runs over its octants in fixed order whatever its neighbours hold: results
// are bitwise reproducible.  Right-hand sides are staged one at a time, which keeps the
// LDS of a block at 17 KB (FP64) for any number of them.
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
    const PcgState *st;    // kApply, kApplyFused, kResid: skip finished right-hand sides
    // kApplyFused: x is formed as z + gamma hold (z on a reset) and written to hnew;
    // with fuse_u, u -= lambda hold (gamma, reset, lambda of the previous iteration)
    const T *z, *hold;
    T *hnew;
    double *u;
    int fuse_u;
};

__device__ __forceinline__ int pmod(int v, int n) { v %= n; return v < 0 ? v + n : v; }

// Component c of neighbour nb = (dz+1)*9 + (dy+1)*3 + (dx+1) relative to the LDS centre hr.
#define E3X_NB(hr, c, nb) (hr)[(c) * HN + ((nb) / 9 - 1) * (HX * HY) + ((nb) / 3 % 3 - 1) * HX + ((nb) % 3 - 1)]

// The body takes the arrays as restrict parameters: the output stores of one right-hand
// side then provably leave the tables untouched, and the coefficient loads of the next
// one stay scalar.
template <typename T, int NR, int MODE>
__device__ __forceinline__ void stencil_body(const StencilArgs<T> &A, const uint16_t *__restrict__ pix,
                                             const T *__restrict__ dk, const T *__restrict__ sten,
                                             const T *__restrict__ x, T *__restrict__ y,
                                             const double *__restrict__ bvec, float *__restrict__ y32,
                                             const T *__restrict__ zv, const T *__restrict__ hold,
                                             T *__restrict__ hnew, double *__restrict__ u)
{
    constexpr bool FUSED = MODE == kApplyFused;
    __shared__ T hs[3 * HN];
    __shared__ int ps[VN];
    __shared__ int cnt[kBlock / 64];
    __shared__ short order[kBlock];
    __shared__ unsigned arrived;  // waves done, for the block partials

    const int bid = xcd_remap(blockIdx.x, gridDim.x);
    const int tx = bid % A.ntx, ty = (bid / A.ntx) % A.nty, tz = bid / (A.ntx * A.nty);
    const int ox = tx * TX, oy = ty * TY, oz = tz * TZ;
    const int nx = A.nx, ny = A.ny, nz = A.nz, nxy = nx * ny;
    const int64_t ns = A.ns;
    const int tid = threadIdx.x;
    if (tid == 0) arrived = 0u;  // ordered before use by the barriers below
    // Interior tiles index without wrapping; boundary tiles wrap each coordinate.
    const bool interior = ox >= 1 && oy >= 1 && oz >= 1 && ox + TX + 1 <= nx && oy + TY + 1 <= ny && oz + TZ + 1 <= nz;
    const int corner = (oz - 1) * nxy + (oy - 1) * nx + (ox - 1);
    unsigned live = (1u << NR) - 1u;  // right-hand sides to compute
    if constexpr (MODE == kApply || MODE == kApplyFused || MODE == kResid) {
        live = 0u;
#pragma unroll
        for (int r = 0; r < NR; r++)
            if (A.st[r].stop_iter == 0) live |= 1u << r;
    }
    auto stage = [&](int r) {
        const int64_t o = (int64_t)(3 * r) * ns;
        if constexpr (FUSED) {
            // this iteration's direction, h = z + gamma h_prev or h = z on a reset
            // (f90:1961-1996); the block's own nodes also store it and take the previous
            // step into u, u -= lambda h_prev (f90:1914-1918)
            const double gam = A.st[r].gamma, lam = A.st[r].lambda;
            const bool rst = A.st[r].reset != 0, run = (live >> r) & 1u, dou = run && A.fuse_u;
            auto put = [&](int q, int g) {
                const int hx = q % HX, hy = (q / HX) % HY, hz = q / (HX * HY);
                const bool own = run && hx >= 1 && hx <= TX && hy >= 1 && hy <= TY && hz >= 1 && hz <= TZ &&
                                 ox + hx - 1 < nx && oy + hy - 1 < ny && oz + hz - 1 < nz;
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    const int64_t i = o + c * ns + g;
                    const T zc = zv[i], hc = hold[i];
                    const T hv = rst ? zc : (T)((double)zc + gam * (double)hc);
                    hs[c * HN + q] = hv;
                    if (own) {
                        hnew[i] = hv;
                        if (dou) u[i] = u[i] - lam * (double)hc;
                    }
                }
            };
            if (interior) {
#pragma unroll
                for (int it = 0; it < (HN + kBlock - 1) / kBlock; it++) {
                    const int q = tid + it * kBlock;
                    if (q < HN) put(q, corner + (q / (HX * HY)) * nxy + ((q / HX) % HY) * nx + q % HX);
                }
            } else {
                for (int q = tid; q < HN; q += kBlock)
                    put(q, pmod(ox + q % HX - 1, nx) + nx * pmod(oy + (q / HX) % HY - 1, ny) +
                               nxy * pmod(oz + q / (HX * HY) - 1, nz));
            }
        } else {
            const T *x0 = x + o, *x1 = x0 + ns, *x2 = x1 + ns;
            if (interior) {  // unrolled: all loads of a thread in flight together
#pragma unroll
                for (int it = 0; it < (HN + kBlock - 1) / kBlock; it++) {
                    const int q = tid + it * kBlock;
                    if (q < HN) {
                        const int g = corner + (q / (HX * HY)) * nxy + ((q / HX) % HY) * nx + q % HX;
                        hs[q] = x0[g]; hs[HN + q] = x1[g]; hs[2 * HN + q] = x2[g];
                    }
                }
            } else {
                for (int q = tid; q < HN; q += kBlock) {
                    const int g = pmod(ox + q % HX - 1, nx) + nx * pmod(oy + (q / HX) % HY - 1, ny) +
                                  nxy * pmod(oz + q / (HX * HY) - 1, nz);
                    hs[q] = x0[g]; hs[HN + q] = x1[g]; hs[2 * HN + q] = x2[g];
                }
            }
        }
    };

    stage(0);  // issued before the early exit below, so its latency overlaps the state loads
    if (interior) {
#pragma unroll
        for (int it = 0; it < (VN + kBlock - 1) / kBlock; it++) {
            const int q = tid + it * kBlock;
            if (q < VN) ps[q] = pix[corner + (q / (VX * VY)) * nxy + ((q / VX) % VY) * nx + q % VX];
        }
    } else {
        for (int q = tid; q < VN; q += kBlock)
            ps[q] = pix[pmod(ox + q % VX - 1, nx) + nx * pmod(oy + (q / VX) % VY - 1, ny) +
                        nxy * pmod(oz + q / (VX * VY) - 1, nz)];
    }
    if (!live) return;  // uniform over the block
    __syncthreads();

    // Stable compaction: homogeneous nodes first, then mixed ones.
    const int w = tid >> 6, l = tid & 63;
    {
        const int lx = (w & 1) * 4 + (l & 3), ly = (w >> 1) * 4 + ((l >> 2) & 3), lz = l >> 4;
        const int *pc = ps + (lz + 1) * (VX * VY) + (ly + 1) * VX + (lx + 1);
        const int p0 = pc[0];
        bool hom = true;
#pragma unroll
        for (int a = 1; a < 8; a++) hom &= pc[-(PZ(a) * (VX * VY) + PY(a) * VX + PX(a))] == p0;
        const unsigned long long bh = __ballot(hom);
        if (l == 0) cnt[w] = __popcll(bh);
        __syncthreads();
        int nh = 0, before = 0;
#pragma unroll
        for (int q = 0; q < kBlock / 64; q++) { nh += cnt[q]; before += q < w ? cnt[q] : 0; }
        const int rh = __popcll(bh & ((1ull << l) - 1ull));
        order[hom ? before + rh : nh + (w * 64 - before) + (l - rh)] = (short)tid;
        __syncthreads();
    }
    const int t = order[tid];
    const int tw = t >> 6, tl = t & 63;
    const int lx = (tw & 1) * 4 + (tl & 3), ly = (tw >> 1) * 4 + ((tl >> 2) & 3), lz = tl >> 4;
    const int gx = ox + lx, gy = oy + ly, gz = oz + lz;
    const bool active = gx < nx && gy < ny && gz < nz;
    const int m = gx + nx * gy + nxy * gz;
    const int *pc = ps + (lz + 1) * (VX * VY) + (ly + 1) * VX + (lx + 1);
    int pv[8];
#pragma unroll
    for (int a = 0; a < 8; a++) pv[a] = pc[-(PZ(a) * (VX * VY) + PY(a) * VX + PX(a))];
    bool homog = true;
#pragma unroll
    for (int a = 1; a < 8; a++) homog &= pv[a] == pv[0];
    const bool hlane = active && homog, mlane = active && !homog;
    const T *hr = hs + (lz + 1) * (HX * HY) + (ly + 1) * HX + (lx + 1);

    double part[MODE == kEnergy ? 2 * NR : NR];
#pragma unroll
    for (int q = 0; q < (MODE == kEnergy ? 2 * NR : NR); q++) part[q] = 0.0;

#pragma unroll
    for (int r = 0; r < NR; r++) {
        if (!((live >> r) & 1u)) continue;
        if (r != 0) {
            __syncthreads();  // every wave is done with the previous right-hand side
            stage(r);
            __syncthreads();
        }
        T acc0 = T(0), acc1 = T(0), acc2 = T(0);
        if (__ballot(hlane)) {
            bool pend = hlane;
            while (true) {
                const unsigned long long pm = __ballot(pend);
                if (pm == 0ull) break;
                const int P = __builtin_amdgcn_readlane(pv[0], __builtin_ctzll(pm));
                const T *S = sten + (int64_t)P * kStenHalf;
                T t0 = T(0), t1 = T(0), t2 = T(0);
#pragma unroll
                for (int k = 0; k < 13; k++) {
#pragma unroll
                    for (int n = 0; n < 3; n++) {  // neighbour +delta
                        const T hp = E3X_NB(hr, n, 14 + k);
                        t0 += S[k * 9 + 0 * 3 + n] * hp;
                        t1 += S[k * 9 + 1 * 3 + n] * hp;
                        t2 += S[k * 9 + 2 * 3 + n] * hp;
                    }
#pragma unroll
                    for (int n = 0; n < 3; n++) {  // neighbour -delta: S[-d][j][n] = S[d][n][j]
                        const T hq = E3X_NB(hr, n, 12 - k);
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
                if (pend && pv[0] == P) {
                    acc0 = t0; acc1 = t1; acc2 = t2;
                    pend = false;
                }
            }
        }
        if (__ballot(mlane)) {
#pragma unroll 1
            for (int a = 0; a < 8; a++) {
                // voxel a of the node spans nodes (PX(b)-PX(a), PY(b)-PY(a), PZ(b)-PZ(a)), b = 0..7
                const T *hb = hr - (PZ(a) * (HX * HY) + PY(a) * HX + PX(a));
                T hv[24];
#pragma unroll
                for (int b = 0; b < 8; b++)
#pragma unroll
                    for (int n = 0; n < 3; n++) hv[b * 3 + n] = hb[n * HN + PZ(b) * (HX * HY) + PY(b) * HX + PX(b)];
                const int pa = pc[-(PZ(a) * (VX * VY) + PY(a) * VX + PX(a))];
                bool pend = mlane;
                while (true) {
                    const unsigned long long pm = __ballot(pend);
                    if (pm == 0ull) break;
                    const int P = __builtin_amdgcn_readlane(pa, __builtin_ctzll(pm));
                    const T *D = dk + (int64_t)P * 576 + a * 72;
                    T t0 = T(0), t1 = T(0), t2 = T(0);
#pragma unroll
                    for (int j = 0; j < 24; j++) {
                        t0 += D[j] * hv[j];
                        t1 += D[24 + j] * hv[j];
                        t2 += D[48 + j] * hv[j];
                    }
                    if (pend && pa == P) {
                        acc0 += t0; acc1 += t1; acc2 += t2;
                        pend = false;
                    }
                }
            }
        }

        if (active) {
            const T acc[3] = {acc0, acc1, acc2};
            if constexpr (MODE == kApply || MODE == kApplyFused) {
                double pr = 0.0;
#pragma unroll
                for (int j = 0; j < 3; j++) {
                    y[(int64_t)(3 * r + j) * ns + m] = acc[j];
                    pr += (double)hr[j * HN] * (double)acc[j];
                }
                part[r] += pr;
            } else if constexpr (MODE == kEnergy) {
                double e = 0.0, gsq = 0.0;
#pragma unroll
                for (int j = 0; j < 3; j++) {
                    const int64_t q = (int64_t)(3 * r + j) * ns + m;
                    const double uq = (double)hr[j * HN], bq = bvec[q], kx = (double)acc[j];
                    e = e + 0.5 * uq * kx + bq * uq;  // f90:1336
                    const double gq = kx + bq;         // f90:1337
                    y[q] = (T)gq;
                    gsq += gq * gq;                     // f90:452
                }
                part[r] += e;
                part[NR + r] += gsq;
            } else if constexpr (MODE == kResid) {
#pragma unroll
                for (int j = 0; j < 3; j++) {
                    const int64_t q = (int64_t)(3 * r + j) * ns + m;
                    y32[q] = (float)((double)acc[j] + bvec[q]);
                }
            } else {
#pragma unroll
                for (int j = 0; j < 3; j++) y[(int64_t)(3 * r + j) * ns + m] = acc[j];
            }
        }
    }
    if constexpr (MODE == kApply || MODE == kApplyFused || MODE == kEnergy) write_partials_nowait(part, A.partials, &arrived);
}
#undef E3X_NB

// Six waves per SIMD: enough to cover the scalar-cache latency of the coefficients,
// within the 80 VGPRs the body needs.
template <typename T, int NR, int MODE>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(6, 8))) void k_stencil(StencilArgs<T> A)
{
    stencil_body<T, NR, MODE>(A, A.pix, A.dk, A.sten, A.x, A.y, A.b, A.y32, A.z, A.hold, A.hnew, A.u);
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

int64_t stencil_parts(const Solver &s) { return stencil_blocks(s); }

hipError_t launch_energy(Solver &s)
{
    StencilArgs<double> A = stencil_args<double>(s, s.dk, s.sten, s.u, s.g);
    A.b = s.b;
    dispatch_stencil<double, kEnergy>(s, A);
    const int np = (int)stencil_parts(s);
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

// Iteration 1 applies K to h as it stands; iteration it >= 2 forms its direction from
// z and the previous one (h_alt / h by parity) while staging.
template <typename T>
void apply_iter(Solver &s, const T *dk, const T *sten, T *hbuf[2], const T *z, T *ah, int it, bool fuse_u)
{
    if (it <= 1) {
        dispatch_stencil<T, kApply>(s, stencil_args<T>(s, dk, sten, hbuf[1], ah));
    } else {
        StencilArgs<T> A = stencil_args<T>(s, dk, sten, nullptr, ah);
        A.z = z; A.hold = hbuf[(it - 1) & 1]; A.hnew = hbuf[it & 1];
        A.u = s.u; A.fuse_u = fuse_u ? 1 : 0;
        dispatch_stencil<T, kApplyFused>(s, A);
    }
}

hipError_t launch_apply_hah(Solver &s, bool mixed, int it, bool fuse_u)
{
    if (mixed) {
        float *hb[2] = {s.h32_alt, s.h32};
        apply_iter<float>(s, s.dk32, s.sten32, hb, s.z32, s.ah32, it, fuse_u);
    } else {
        double *hb[2] = {s.h_alt, s.h};
        apply_iter<double>(s, s.dk, s.sten, hb, s.z, s.ah, it, fuse_u);
    }
    return launch_hah_finish(s, (int)stencil_parts(s), it);
}

// The fused apply of an even iteration, as the PCG loop runs it (u and the state are
// left to the caller: with lambda = gamma = 0 it leaves u unchanged and sets h = z).
hipError_t launch_apply_only(Solver &s, bool mixed)
{
    if (mixed) {
        float *hb[2] = {s.h32_alt, s.h32};
        apply_iter<float>(s, s.dk32, s.sten32, hb, s.z32, s.ah32, 2, true);
    } else {
        double *hb[2] = {s.h_alt, s.h};
        apply_iter<double>(s, s.dk, s.sten, hb, s.z, s.ah, 2, true);
    }
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
