// gfx950 matrix-free stencil y = K x of the elas3dxtal_pcg solve (energy_OpenMP
// f90:1176-1321, dembx SpMV f90:1585-1884).
//
// A block owns an 8x8x4 tile of nodes and stages the 10x10x6 node halo of x and the
// 9x9x5 voxel phases in LDS.  A node is homogeneous when its 8 voxels share one phase,
// mixed otherwise.  The lanes take the tile's nodes in an order fixed when the
// microstructure is loaded (k_tile_order): homogeneous nodes first, by phase, then the
// mixed ones by their phases, so that most waves hold one kind and few phases.  Coefficients are wave-uniform
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
// are bitwise reproducible.  Each block serves one right-hand side, so the
// kernel is the same for any number of them (17 KB of LDS in FP64).
#include "e3x_device.hpp"

#include <cmath>

namespace e3x {
namespace {

template <typename T>
struct StencilArgs {
    int nx, ny, nz, ntx, nty, nrhs;
    int64_t ns;
    const uint16_t *pix;
    const uint8_t *perm;   // lane order per tile (Solver::perm)
    const T *dk, *sten;
    const T *x;            // [nrhs][3][ns]
    T *y;                  // [nrhs][3][ns] (kApply, kEnergy, kPlain)
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

// Wave w of a block owns the 4x4x4 cube at (wave_x, wave_y, wave_z) of the tile.
__device__ __forceinline__ int wave_x(int w) { return (w % (TX / 4)) * 4; }
__device__ __forceinline__ int wave_y(int w) { return ((w / (TX / 4)) % (TY / 4)) * 4; }
__device__ __forceinline__ int wave_z(int w) { return (w / ((TX / 4) * (TY / 4))) * 4; }

// The phase tables are read-only while a stencil kernel runs.  Reading them through the
// constant address space makes every wave-uniform coefficient load a scalar load
// whatever the kernel stores elsewhere (through the global address space the compiler
// must prove that no store clobbers them, which conditional stores defeat).
template <typename T>
using ctab = const __attribute__((address_space(4))) T;
template <typename T>
__device__ __forceinline__ ctab<T> *as_table(const T *p) { return (ctab<T> *)p; }

// Component c of neighbour nb = (dz+1)*9 + (dy+1)*3 + (dx+1) relative to the LDS centre hr.
#define E3X_NB(hr, c, nb) (hr)[(c) * HN + ((nb) / 9 - 1) * (HX * HY) + ((nb) / 3 % 3 - 1) * HX + ((nb) % 3 - 1)]

// The body takes the arrays as restrict parameters: the output stores then provably
// leave the tables untouched, and coefficient loads after a store stay scalar.
//
// RB right-hand sides per block (grid: tiles x ceil(nrhs / RB)): the block stages the RB
// halos, classifies its nodes once and applies every scalar coefficient it loads to all
// RB of them.  A right-hand side past nrhs, or one that has stopped, is computed from
// whatever its LDS slots hold and not stored.
template <typename T, int MODE, int RB>
__device__ __forceinline__ void stencil_body(const StencilArgs<T> &A, const uint16_t *__restrict__ pix,
                                             const T *__restrict__ dk, const T *__restrict__ sten,
                                             const T *__restrict__ x, T *__restrict__ y,
                                             const double *__restrict__ bvec, float *__restrict__ y32,
                                             const T *__restrict__ zv, const T *__restrict__ hold,
                                             T *__restrict__ hnew, double *__restrict__ u)
{
    constexpr bool FUSED = MODE == kApplyFused;
    constexpr bool HAS_ST = MODE == kApply || MODE == kApplyFused || MODE == kResid;
    constexpr int NP = MODE == kEnergy ? 2 : 1;  // partial sums per right-hand side
    __shared__ T hs[RB * 3 * HN];
    __shared__ int ps[VN];
    __shared__ unsigned arrived;  // waves done, for the block partials

    // grid (tiles, groups of RB right-hand sides); the XCD remap runs over the linear
    // block id, whose remainder mod 8 is the XCD, and (tile, group) come from the
    // remapped id
    const int ntile = gridDim.x, ngroup = gridDim.y, nrhs = A.nrhs;
    const int lin = xcd_remap(blockIdx.x + blockIdx.y * ntile, ntile * ngroup);
    const int grp = lin / ntile, bid = lin % ntile;
    const int tx = bid % A.ntx, ty = (bid / A.ntx) % A.nty, tz = bid / (A.ntx * A.nty);
    const int ox = tx * TX, oy = ty * TY, oz = tz * TZ;
    const int nx = A.nx, ny = A.ny, nz = A.nz, nxy = nx * ny;
    const int64_t ns = A.ns;
    const int tid = threadIdx.x;
    if (tid == 0) arrived = 0u;  // ordered before use by the barrier below
    // this lane's node: t in the tile's wave-cube numbering (homogeneous nodes by phase,
    // then mixed ones by their phases; k_tile_order)
    const int t = A.perm[(int64_t)bid * kSB + tid];
    // Interior tiles index without wrapping; boundary tiles wrap each coordinate.
    const bool interior = ox >= 1 && oy >= 1 && oz >= 1 && ox + TX + 1 <= nx && oy + TY + 1 <= ny && oz + TZ + 1 <= nz;
    const int corner = (oz - 1) * nxy + (oy - 1) * nx + (ox - 1);

    unsigned live = 0u;  // bit rb: right-hand side grp*RB + rb exists and still runs
#pragma unroll
    for (int rb = 0; rb < RB; rb++) {
        const int r = grp * RB + rb;
        if (r < nrhs && (!HAS_ST || A.st[r].stop_iter == 0)) live |= 1u << rb;
    }
    auto stage = [&](int rb) {
        const int r = grp * RB + rb;
        const int64_t o = (int64_t)(3 * r) * ns;
        T *hsr = hs + rb * 3 * HN;
        if constexpr (FUSED) {
            // this iteration's direction, h = z + gamma h_prev or h = z on a reset
            // (f90:1961-1996); the block's own nodes also store it and take the previous
            // step into u, u -= lambda h_prev (f90:1914-1918)
            const double gam = A.st[r].gamma, lam = A.st[r].lambda;
            const bool run = (live >> rb) & 1u, rst = A.st[r].reset != 0, dou = A.fuse_u != 0;
            auto put = [&](int q, int g) {
                const int hx = q % HX, hy = (q / HX) % HY, hz = q / (HX * HY);
                const bool own = run && hx >= 1 && hx <= TX && hy >= 1 && hy <= TY && hz >= 1 && hz <= TZ &&
                                 ox + hx - 1 < nx && oy + hy - 1 < ny && oz + hz - 1 < nz;
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    const int64_t i = o + c * ns + g;
                    const T zc = zv[i], hc = hold[i];
                    const T hv = rst ? zc : (T)((double)zc + gam * (double)hc);
                    hsr[c * HN + q] = hv;
                    if (own) {
                        hnew[i] = hv;
                        if (dou) u[i] = u[i] - lam * (double)hc;
                    }
                }
            };
            if (interior) {
#pragma unroll
                for (int it = 0; it < (HN + kSB - 1) / kSB; it++) {
                    const int q = tid + it * kSB;
                    if (q < HN) put(q, corner + (q / (HX * HY)) * nxy + ((q / HX) % HY) * nx + q % HX);
                }
            } else {
                for (int q = tid; q < HN; q += kSB)
                    put(q, pmod(ox + q % HX - 1, nx) + nx * pmod(oy + (q / HX) % HY - 1, ny) +
                               nxy * pmod(oz + q / (HX * HY) - 1, nz));
            }
        } else {
            const T *x0 = x + o, *x1 = x0 + ns, *x2 = x1 + ns;
            if (interior) {  // unrolled: all loads of a thread in flight together
#pragma unroll
                for (int it = 0; it < (HN + kSB - 1) / kSB; it++) {
                    const int q = tid + it * kSB;
                    if (q < HN) {
                        const int g = corner + (q / (HX * HY)) * nxy + ((q / HX) % HY) * nx + q % HX;
                        hsr[q] = x0[g]; hsr[HN + q] = x1[g]; hsr[2 * HN + q] = x2[g];
                    }
                }
            } else {
                for (int q = tid; q < HN; q += kSB) {
                    const int g = pmod(ox + q % HX - 1, nx) + nx * pmod(oy + (q / HX) % HY - 1, ny) +
                                  nxy * pmod(oz + q / (HX * HY) - 1, nz);
                    hsr[q] = x0[g]; hsr[HN + q] = x1[g]; hsr[2 * HN + q] = x2[g];
                }
            }
        }
    };

    // the voxel phases first: their loads overlap the state loads of the early exit
    if (interior) {
#pragma unroll
        for (int it = 0; it < (VN + kSB - 1) / kSB; it++) {
            const int q = tid + it * kSB;
            if (q < VN) ps[q] = pix[corner + (q / (VX * VY)) * nxy + ((q / VX) % VY) * nx + q % VX];
        }
    } else {
        for (int q = tid; q < VN; q += kSB)
            ps[q] = pix[pmod(ox + q % VX - 1, nx) + nx * pmod(oy + (q / VX) % VY - 1, ny) +
                        nxy * pmod(oz + q / (VX * VY) - 1, nz)];
    }
    // One staging site only: with a second one (staging before this exit) the compiler
    // no longer proves the tables unclobbered and loads the coefficients per lane.
    if (!live) return;  // uniform over the block
#pragma unroll
    for (int rb = 0; rb < RB; rb++)
        if ((live >> rb) & 1u) stage(rb);
    __syncthreads();

    const int tw = t >> 6, tl = t & 63;
    const int lx = wave_x(tw) + (tl & 3), ly = wave_y(tw) + ((tl >> 2) & 3), lz = wave_z(tw) + (tl >> 4);
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
    const T *hr = hs + (lz + 1) * (HX * HY) + (ly + 1) * HX + (lx + 1);  // rhs rb: hr + rb * 3 * HN

    T acc[RB][3];
#pragma unroll
    for (int rb = 0; rb < RB; rb++) acc[rb][0] = acc[rb][1] = acc[rb][2] = T(0);
    if (__ballot(hlane)) {
        bool pend = hlane;
        while (true) {
            const unsigned long long pm = __ballot(pend);
            if (pm == 0ull) break;
            const int P = __builtin_amdgcn_readlane(pv[0], __builtin_ctzll(pm));
            ctab<T> *S = as_table(sten) + (int64_t)P * kStenHalf;
            T tt[RB][3];
#pragma unroll
            for (int rb = 0; rb < RB; rb++) tt[rb][0] = tt[rb][1] = tt[rb][2] = T(0);
#pragma unroll
            for (int k = 0; k < 13; k++) {
#pragma unroll
                for (int n = 0; n < 3; n++) {  // neighbour +delta
                    const T s0 = S[k * 9 + 0 * 3 + n], s1 = S[k * 9 + 1 * 3 + n], s2 = S[k * 9 + 2 * 3 + n];
#pragma unroll
                    for (int rb = 0; rb < RB; rb++) {
                        const T hp = E3X_NB(hr + rb * 3 * HN, n, 14 + k);
                        tt[rb][0] += s0 * hp; tt[rb][1] += s1 * hp; tt[rb][2] += s2 * hp;
                    }
                }
#pragma unroll
                for (int n = 0; n < 3; n++) {  // neighbour -delta: S[-d][j][n] = S[d][n][j]
                    const T s0 = S[k * 9 + n * 3 + 0], s1 = S[k * 9 + n * 3 + 1], s2 = S[k * 9 + n * 3 + 2];
#pragma unroll
                    for (int rb = 0; rb < RB; rb++) {
                        const T hq = E3X_NB(hr + rb * 3 * HN, n, 12 - k);
                        tt[rb][0] += s0 * hq; tt[rb][1] += s1 * hq; tt[rb][2] += s2 * hq;
                    }
                }
            }
            ctab<T> *Cc = S + 117;  // centre: 00 11 22 01 02 12
#pragma unroll
            for (int rb = 0; rb < RB; rb++) {
                const T *h0 = hr + rb * 3 * HN;
                const T c0 = h0[0], c1 = h0[HN], c2 = h0[2 * HN];
                tt[rb][0] += Cc[0] * c0; tt[rb][0] += Cc[3] * c1; tt[rb][0] += Cc[4] * c2;
                tt[rb][1] += Cc[3] * c0; tt[rb][1] += Cc[1] * c1; tt[rb][1] += Cc[5] * c2;
                tt[rb][2] += Cc[4] * c0; tt[rb][2] += Cc[5] * c1; tt[rb][2] += Cc[2] * c2;
            }
            if (pend && pv[0] == P) {
#pragma unroll
                for (int rb = 0; rb < RB; rb++) {
                    acc[rb][0] = tt[rb][0]; acc[rb][1] = tt[rb][1]; acc[rb][2] = tt[rb][2];
                }
                pend = false;
            }
        }
    }
    if (__ballot(mlane)) {
#pragma unroll 1
        for (int a = 0; a < 8; a++) {
            // voxel a of the node spans nodes (PX(b)-PX(a), PY(b)-PY(a), PZ(b)-PZ(a)), b = 0..7
            const T *hb = hr - (PZ(a) * (HX * HY) + PY(a) * HX + PX(a));
            T hv[RB][24];
#pragma unroll
            for (int rb = 0; rb < RB; rb++)
#pragma unroll
                for (int b = 0; b < 8; b++)
#pragma unroll
                    for (int n = 0; n < 3; n++)
                        hv[rb][b * 3 + n] = hb[rb * 3 * HN + n * HN + PZ(b) * (HX * HY) + PY(b) * HX + PX(b)];
            const int pa = pc[-(PZ(a) * (VX * VY) + PY(a) * VX + PX(a))];
            bool pend = mlane;
            while (true) {
                const unsigned long long pm = __ballot(pend);
                if (pm == 0ull) break;
                const int P = __builtin_amdgcn_readlane(pa, __builtin_ctzll(pm));
                ctab<T> *D = as_table(dk) + (int64_t)P * 576 + a * 72;
                T tt[RB][3];
#pragma unroll
                for (int rb = 0; rb < RB; rb++) tt[rb][0] = tt[rb][1] = tt[rb][2] = T(0);
#pragma unroll
                for (int j = 0; j < 24; j++) {
                    const T d0 = D[j], d1 = D[24 + j], d2 = D[48 + j];
#pragma unroll
                    for (int rb = 0; rb < RB; rb++) {
                        tt[rb][0] += d0 * hv[rb][j]; tt[rb][1] += d1 * hv[rb][j]; tt[rb][2] += d2 * hv[rb][j];
                    }
                }
                if (pend && pa == P) {
#pragma unroll
                    for (int rb = 0; rb < RB; rb++) {
                        acc[rb][0] += tt[rb][0]; acc[rb][1] += tt[rb][1]; acc[rb][2] += tt[rb][2];
                    }
                    pend = false;
                }
            }
        }
    }

    double part[RB * NP];
#pragma unroll
    for (int q = 0; q < RB * NP; q++) part[q] = 0.0;
#pragma unroll
    for (int rb = 0; rb < RB; rb++) {
        if (!active || !((live >> rb) & 1u)) continue;
        const int64_t ro = (int64_t)(3 * (grp * RB + rb)) * ns;  // this right-hand side's vectors
        const T *h0 = hr + rb * 3 * HN;
        if constexpr (MODE == kApply || MODE == kApplyFused) {
            double pr = 0.0;
#pragma unroll
            for (int j = 0; j < 3; j++) {
                y[ro + j * ns + m] = acc[rb][j];
                pr += (double)h0[j * HN] * (double)acc[rb][j];
            }
            part[rb] = pr;
        } else if constexpr (MODE == kEnergy) {
            double e = 0.0, gsq = 0.0;
#pragma unroll
            for (int j = 0; j < 3; j++) {
                const int64_t q = ro + j * ns + m;
                const double uq = (double)h0[j * HN], bq = bvec[q], kx = (double)acc[rb][j];
                e = e + 0.5 * uq * kx + bq * uq;  // f90:1336
                const double gq = kx + bq;         // f90:1337
                y[q] = (T)gq;
                gsq += gq * gq;                     // f90:452
            }
            part[rb] = e;
            part[RB + rb] = gsq;
        } else if constexpr (MODE == kResid) {
#pragma unroll
            for (int j = 0; j < 3; j++) {
                const int64_t q = ro + j * ns + m;
                y32[q] = (float)((double)acc[rb][j] + bvec[q]);
            }
        } else {
#pragma unroll
            for (int j = 0; j < 3; j++) y[ro + j * ns + m] = acc[rb][j];
        }
    }
    // partial slots [q * nrhs + r][tile]: h.Ah (or utot), then gg for kEnergy; none for a
    // right-hand side past nrhs
    if constexpr (MODE == kApply || MODE == kApplyFused || MODE == kEnergy) {
        int64_t slot[RB * NP];
#pragma unroll
        for (int rb = 0; rb < RB; rb++) {
            const int r = grp * RB + rb;
            slot[rb] = r < nrhs ? (int64_t)r * ntile + bid : -1;
            if constexpr (MODE == kEnergy) slot[RB + rb] = r < nrhs ? (int64_t)(nrhs + r) * ntile + bid : -1;
        }
        write_partials_nowait<RB * NP, kSB>(part, A.partials, slot, &arrived);
    }
}
#undef E3X_NB

// At least WPE waves per SIMD: six for the FP64 body (80 VGPRs; forced into 64 VGPRs
// for eight waves it hoists fewer loads and ran 7 % slower), eight for the FP32 body.
template <typename T, int MODE, int RB, int WPE>
__global__ __launch_bounds__(kSB) __attribute__((amdgpu_waves_per_eu(WPE, 8))) void k_stencil(StencilArgs<T> A)
{
    stencil_body<T, MODE, RB>(A, A.pix, A.dk, A.sten, A.x, A.y, A.b, A.y32, A.z, A.hold, A.hnew, A.u);
}

// Lane order of each stencil tile (Solver::perm), from the phases alone: one block per
// tile, thread t = node t of the tile's wave-cube numbering.  Key: homogeneous nodes
// first, by phase; then mixed nodes by (smallest, largest, octant-0 phase); nodes
// outside the grid last; ties by t.  A bitonic sort of the 256 keys in LDS.
__global__ __launch_bounds__(kSB) void k_tile_order(int nx, int ny, int nz, int ntx, int nty, const uint16_t *pix,
                                                  uint8_t *perm)
{
    __shared__ int ps[VN];
    __shared__ unsigned long long key[kSB];
    const int bid = blockIdx.x, t = threadIdx.x;
    const int tx = bid % ntx, ty = (bid / ntx) % nty, tz = bid / (ntx * nty);
    const int ox = tx * TX, oy = ty * TY, oz = tz * TZ, nxy = nx * ny;
    for (int q = t; q < VN; q += kSB)
        ps[q] = pix[pmod(ox + q % VX - 1, nx) + nx * pmod(oy + (q / VX) % VY - 1, ny) +
                    nxy * pmod(oz + q / (VX * VY) - 1, nz)];
    __syncthreads();
    const int tw = t >> 6, tl = t & 63;
    const int lx = wave_x(tw) + (tl & 3), ly = wave_y(tw) + ((tl >> 2) & 3), lz = wave_z(tw) + (tl >> 4);
    const int *pc = ps + (lz + 1) * (VX * VY) + (ly + 1) * VX + (lx + 1);
    unsigned pmin = 0xFFFFu, pmax = 0u;
    const unsigned p0 = (unsigned)pc[0];
#pragma unroll
    for (int a = 0; a < 8; a++) {
        const unsigned pa = (unsigned)pc[-(PZ(a) * (VX * VY) + PY(a) * VX + PX(a))];
        pmin = min(pmin, pa);
        pmax = max(pmax, pa);
    }
    const bool active = ox + lx < nx && oy + ly < ny && oz + lz < nz;
    const unsigned long long cls = !active ? 2ull : (pmin == pmax ? 0ull : 1ull);
    key[t] = (cls << 62) | ((unsigned long long)pmin << 44) | ((unsigned long long)pmax << 28) |
             ((unsigned long long)p0 << 12) | (unsigned long long)t;
    __syncthreads();
    for (int k = 2; k <= kSB; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            const int i = t, p = t ^ j;
            if (p > i) {
                const unsigned long long a = key[i], b = key[p];
                if (((i & k) == 0) == (a > b)) {
                    key[i] = b;
                    key[p] = a;
                }
            }
            __syncthreads();
        }
    perm[(int64_t)bid * kSB + t] = (uint8_t)(key[t] & 0xFFull);
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
    A.nx = s.nx; A.ny = s.ny; A.nz = s.nz; A.nrhs = s.nrhs;
    A.ntx = (s.nx + TX - 1) / TX; A.nty = (s.ny + TY - 1) / TY;
    A.ns = s.ns; A.pix = s.pix; A.perm = s.perm; A.dk = dk; A.sten = sten; A.x = x; A.y = y;
    A.partials = s.partials; A.st = s.st;
    return A;
}

// One right-hand side per block (grid y = nrhs).  The body also serves RB > 1 (shared
// staging of the phases, classification and coefficient loads), but at 200^3 two FP32
// right-hand sides per block ran 2-4 % slower than one: the per-rhs FMA, LDS and
// staging work dominates, and the second halo costs occupancy.
template <typename T, int MODE>
void dispatch_stencil(const Solver &s, const StencilArgs<T> &A)
{
    const dim3 grid((unsigned)stencil_blocks(s), (unsigned)s.nrhs), block(kSB);
    if constexpr (sizeof(T) == 4) hipLaunchKernelGGL((k_stencil<T, MODE, 1, 8>), grid, block, 0, s.stream, A);
    else hipLaunchKernelGGL((k_stencil<T, MODE, 1, 6>), grid, block, 0, s.stream, A);
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

hipError_t launch_tile_order(Solver &s)
{
    hipLaunchKernelGGL(k_tile_order, dim3((unsigned)stencil_blocks(s)), dim3(kSB), 0, s.stream, s.nx, s.ny, s.nz,
                       (s.nx + TX - 1) / TX, (s.ny + TY - 1) / TY, (const uint16_t *)s.pix, s.perm);
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
