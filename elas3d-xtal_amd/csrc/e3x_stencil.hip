// gfx950 matrix-free stencil y = K x of the elas3dxtal_pcg solve (energy_OpenMP
// f90:1176-1321, dembx SpMV f90:1585-1884).
//
// A block owns an 8x8x4 tile of nodes and stages the 10x10x6 node halo of x and the
// 9x9x5 voxel phases in LDS.  A node is homogeneous when its 8 voxels share one phase,
// mixed otherwise.  The lanes take the tile's nodes in an order fixed when the
// microstructure is loaded (k_tile_order): homogeneous nodes first, by phase, then the
// mixed ones by their phases, so that most waves hold one kind and few phases.
// Coefficients are wave-uniform and read through the scalar cache (SGPR operands, no
// per-lane coefficient loads):
//   - homogeneous lanes: the wave loops over the phases its lanes hold; for phase P the
//     27-point stencil of P (243 FMA; the block of -delta is the transpose of the block
//     of +delta, so 126 coefficients) goes to the lanes of phase P;
//   - mixed nodes: shared by the block's four waves, wave w taking the octant pair
//     (w, w^6) of opposite corners for every mixed node of the tile: per octant a lane
//     reads the 24 neighbour values of that voxel, the wave loops over the phases its
//     lanes have there and applies 3x24 rows of the element matrix (72 FMA) to those
//     lanes; both octants of a pair use the rows dk[P][w], the second in permuted order.
//     The lane holding the node adds the four pair sums in wave order.
// The next phase is the one of the first lane still pending (ballot + readlane).  Each
// lane is served exactly once per octant (or once in all, when homogeneous), so a
// node's sum runs over its octants in fixed order whatever its neighbours hold: results
// are bitwise reproducible.  Each block serves one right-hand side, so the
// kernel is the same for any number of them (24 KB of LDS in FP64).
#include "e3x_device.hpp"

#include <cmath>

namespace e3x {
namespace {

// n / d for 0 <= n < 2^31 by a multiply and a shift (d fixed per launch): with
// 2^s <= d < 2^(s+1) and M = ceil(2^(32+s) / d) <= 2^32, the error M d - 2^(32+s) < d
// keeps floor(n M / 2^(32+s)) exact.  Uniform operands stay in scalar registers (the
// hardware has no integer divide; a division is otherwise a vector reciprocal sequence).
struct FastDiv {
    unsigned long long m = 1;
    int sh = 32;
};
inline FastDiv fast_div(unsigned d)
{
    int s = 0;
    while ((2ull << s) <= d) s++;
    FastDiv f;
    f.m = ((1ull << (32 + s)) + d - 1) / d;
    f.sh = 32 + s;
    return f;
}
__device__ __forceinline__ int fdiv(int n, FastDiv f) { return (int)(((unsigned long long)(unsigned)n * f.m) >> f.sh); }

template <typename T>
struct StencilArgs {
    int nx, ny, nz, ntx, nty, nrhs;
    FastDiv d_ntile, d_ntx, d_nty;  // by the tiles of one right-hand side, tiles along x, along y
    int wrap1;                      // every extent >= 8: a halo index wraps at most once
    int64_t ns;
    const void *pix;       // phase ids: uint16, or uint32 on a wide handle (the kernel's PhT)
    const PermT *perm;     // lane order per tile (Solver::perm)
    const unsigned *tcnt;  // homogeneous and mixed node counts per tile (Solver::tcnt)
    int64_t mstride;       // batch: nodes between the members' phase maps (0: one shared map)
    const T *dk, *sten;
    const T *x;            // [nrhs][ns][3]
    T *y;                  // [nrhs][ns][3] (kApply, kEnergy, kPlain)
    const double *b;       // kEnergy, kResid
    float *y32;            // kResid output
    double *partials;
    const PcgState *st;    // kApply, kApplyFused, kResid: skip finished right-hand sides
    // kApplyFused: x is formed as z + gamma hold (z on a reset) and written to hnew;
    // with usteps = 2, u -= lambda_prev h_{it-2} (read from hnew before it is overwritten),
    // then u -= lambda hold (gamma, reset, lambdas of the previous iterations)
    const T *z, *hold;
    T *hnew;
    ustep_t<T> *u;         // u (FP64), or the mixed form's FP32 u steps since the last replacement
    int usteps;
    int check_first;       // read the stop state before loading (launches past the predicted stop)
    int rmap[kMaxRhs];     // grid row j serves right-hand side rmap[j]
};

__device__ __forceinline__ int pmod(int v, int n) { v %= n; return v < 0 ? v + n : v; }

// Wave w of a block owns the 4x4x4 cube at (wave_x, wave_y, wave_z) of the tile.
__device__ __forceinline__ int wave_x(int w) { return (w % (TX / 4)) * 4; }
__device__ __forceinline__ int wave_y(int w) { return ((w / (TX / 4)) % (TY / 4)) * 4; }
__device__ __forceinline__ int wave_z(int w) { return (w / ((TX / 4) * (TY / 4))) * 4; }

// Halo positions (hx, hy, hz), 0 <= h < H: thread n stages the tile's own node n (x
// fastest, so its stores of the node go out row by row) and the shell positions n and
// n + kSB of the kShell halo nodes around the tile (the two z faces, then per inner
// plane its y = 0 and y = HY-1 rows and its x = 0 and x = HX-1 ends).
struct HPos {
    int x, y, z;
};
constexpr int kShell = HN - kSB;
constexpr int kShellPer = (kShell + kSB - 1) / kSB;  // shell positions per thread
constexpr int kVoxPer = (VN + kSB - 1) / kSB;        // voxel-halo positions per thread
__device__ __forceinline__ HPos inner_pos(int n) { return {1 + n % TX, 1 + (n / TX) % TY, 1 + n / (TX * TY)}; }
__device__ __forceinline__ HPos shell_pos(int s)
{
    constexpr int plane = HX * HY, ring = 2 * HX + 2 * (HY - 2);
    if (s < 2 * plane) {
        const int w = s % plane;
        return {w % HX, w / HX, s < plane ? 0 : HZ - 1};
    }
    const int r = s - 2 * plane, w = r % ring, hz = 1 + r / ring;
    if (w < 2 * HX) return {w % HX, w < HX ? 0 : HY - 1, hz};
    const int v = w - 2 * HX;
    return {(v & 1) * (HX - 1), 1 + (v >> 1), hz};
}
__device__ __forceinline__ int hpos_index(HPos p) { return p.z * (HX * HY) + p.y * HX + p.x; }

// The phase tables are read-only while a stencil kernel runs.  Reading them through the
// constant address space makes every wave-uniform coefficient load a scalar load
// whatever the kernel stores elsewhere (through the global address space the compiler
// must prove that no store clobbers them, which conditional stores defeat).
template <typename T>
using ctab = const __attribute__((address_space(4))) T;
template <typename T>
__device__ __forceinline__ ctab<T> *as_table(const T *p) { return (ctab<T> *)p; }

// Scalar-cache prefetch of a coefficient slice of BYTES bytes: one dword every 64 bytes
// (the line size; the last dword read lies inside the slice),
// all issued before a single wait, so that the slice's cache misses overlap in one round
// trip.  Without it the misses come one group at a time: the compiler splits a slice into
// several scalar loads, each waited for with lgkmcnt(0), which also drains the LDS reads.
// The empty asm consumes the folded value, which keeps the loads.  Used by the FP64 body,
// whose slices are twice the bytes of the FP32 ones (the FP32 body gains nothing from it).
template <int BYTES>
__device__ __forceinline__ void touch_lines(const void *p)
{
    constexpr int N = (BYTES - 4) / 64 + 1;
    const __attribute__((address_space(4))) unsigned *w = (const __attribute__((address_space(4))) unsigned *)p;
    unsigned v[N];
#pragma unroll
    for (int i = 0; i < N; i++) v[i] = w[i * 16];
    unsigned x = 0;
#pragma unroll
    for (int i = 0; i < N; i++) x ^= v[i];
    asm volatile("" ::"s"(x));
}

// Component c of neighbour nb = (dz+1)*9 + (dy+1)*3 + (dx+1) relative to the LDS centre hr.
#define E3X_NB(hr, c, nb) (hr)[(c) * HN + ((nb) / 9 - 1) * (HX * HY) + ((nb) / 3 % 3 - 1) * HX + ((nb) % 3 - 1)]

// The body takes the arrays as restrict parameters: the output stores then provably
// leave the tables untouched, and coefficient loads after a store stay scalar.
//
// One right-hand side per block (grid: tiles x nrhs).  Every global load of the
// staging (phases, this tile's lane order, the halo and, fused, u at the tile's own
// nodes) is issued at the top, before the block learns from the solver state whether
// its right-hand side still runs, so that the state's latency and the loads' overlap;
// a stopped right-hand side drops what it loaded and leaves.
template <typename T, int MODE, typename PhT>
__device__ __forceinline__ void stencil_body(const StencilArgs<T> &A, const PhT *__restrict__ pix,
                                             const T *__restrict__ dk, const T *__restrict__ sten,
                                             const T *__restrict__ x, T *__restrict__ y,
                                             const double *__restrict__ bvec, float *__restrict__ y32,
                                             const T *__restrict__ zv, const T *__restrict__ hold,
                                             T *__restrict__ hnew, ustep_t<T> *__restrict__ u)
{
    constexpr bool FUSED = MODE == kApplyFused;
    constexpr bool HAS_ST = MODE == kApply || MODE == kApplyFused || MODE == kResid;
    constexpr bool DIRECT = sizeof(T) == 8 && (MODE == kApply || MODE == kApplyFused);
    __shared__ T hs[3 * HN];
    __shared__ T as[3 * kSB];  // Ah in x-fastest tile order
    __shared__ int ps[VN];
    __shared__ PermT tps[kSB];   // the lane order (node of each lane)
    __shared__ T part[4][3][64];  // mixed nodes: each wave's partial sums
    __shared__ unsigned arrived;  // waves done (last_wave)

    // grid (tiles, right-hand sides); the XCD remap runs over the linear block id, whose
    // remainder mod 8 is the XCD, and (tile, right-hand side) come from the remapped id.
    // Tiles run from the end of the grid to its start: the streaming update that precedes
    // the apply in the PCG loop writes z from the start to the end, so the first tiles
    // find the most recent z lines still in the last-level cache (-4 % at 200^3 FP64).
    const int ntile = gridDim.x, nrhs = A.nrhs;
    const int lin = xcd_remap(blockIdx.x + blockIdx.y * ntile, ntile * gridDim.y);
    const int row = fdiv(lin, A.d_ntile), bid = ntile - 1 - (lin - row * ntile);
    const int r = A.rmap[row];
    const int txy = fdiv(bid, A.d_ntx), tz = fdiv(txy, A.d_nty);
    const int tx = bid - txy * A.ntx, ty = txy - tz * A.nty;
    const int ox = tx * TX, oy = ty * TY, oz = tz * TZ;
    const int nx = A.nx, ny = A.ny, nz = A.nz, nxy = nx * ny;
    const int64_t ns = A.ns, o = (int64_t)(3 * r) * ns;
    // a batch's member r has its own phase map and lane order
    const int64_t mbase = (int64_t)r * A.mstride, ptile = A.mstride != 0 ? (int64_t)r * ntile * kSB : 0;
    const int tid = threadIdx.x;
    if constexpr (HAS_ST) {
        // a launch the host expects to find its right-hand side stopped checks before it loads
        if (A.check_first != 0 && A.st[r].stop_iter != 0) return;  // uniform over the block
    }
    if (tid == 0) arrived = 0u;  // ordered before use by the barrier below
    // the solver state first, so that its latency overlaps the loads below (r < nrhs:
    // the grid has nrhs rows)
    int stop = 0, rst = 0;
    double gam = 0.0, lam = 0.0, lamp = 0.0;
    if constexpr (HAS_ST) stop = A.st[r].stop_iter;
    if constexpr (FUSED) {
        gam = A.st[r].gamma;
        lam = A.st[r].lambda;
        lamp = A.st[r].lambda_prev;
        rst = A.st[r].reset;
    }
    // Interior tiles index without wrapping; boundary tiles wrap each coordinate.
    const bool interior = ox >= 1 && oy >= 1 && oz >= 1 && ox + TX + 1 <= nx && oy + TY + 1 <= ny && oz + TZ + 1 <= nz;
    const int corner = (oz - 1) * nxy + (oy - 1) * nx + (ox - 1);
    auto gpos = [&](HPos p) -> int {
        return interior ? corner + p.z * nxy + p.y * nx + p.x
                        : A.wrap1 ? wrap(ox + p.x - 1, nx) + nx * wrap(oy + p.y - 1, ny) + nxy * wrap(oz + p.z - 1, nz)
                                  : pmod(ox + p.x - 1, nx) + nx * pmod(oy + p.y - 1, ny) + nxy * pmod(oz + p.z - 1, nz);
    };

    // --- loads -------------------------------------------------------------------
    // voxel phases of the 9x9x5 voxel halo (positions tid and tid + kSB)
    int pvl[kVoxPer];
#pragma unroll
    for (int k = 0; k < kVoxPer; k++) {
        const int q = min(tid + k * kSB, VN - 1);
        const HPos p = {q % VX, (q / VX) % VY, q / (VX * VY)};
        pvl[k] = (int)pix[mbase + gpos(p)];
    }
    // this lane's node: t in the tile's wave-cube numbering (homogeneous nodes by phase,
    // then mixed ones by their phases; k_tile_order)
    const int t = A.perm[ptile + (int64_t)bid * kSB + tid];
    // halo: the tile's own node tid and shell positions tid + k kSB (a thread without
    // one loads its own node again and does not store it)
    const HPos pin = inner_pos(tid);
    HPos shp[kShellPer];
    bool shv[kShellPer];
    int64_t gs[kShellPer];
#pragma unroll
    for (int k = 0; k < kShellPer; k++) {
        shv[k] = tid + k * kSB < kShell;
        shp[k] = shv[k] ? shell_pos(tid + k * kSB) : pin;
        gs[k] = o + 3 * (int64_t)gpos(shp[k]);
    }
    const int64_t gi = o + 3 * (int64_t)gpos(pin);
    const bool ingrid = ox + pin.x - 1 < nx && oy + pin.y - 1 < ny && oz + pin.z - 1 < nz;
    const T *src = FUSED ? zv : x;
    T ai[3], as_[kShellPer][3];  // z (fused) or x
    T bi[3], bs_[kShellPer][3];  // h_prev (fused)
    double ui[3] = {0.0, 0.0, 0.0};
    T hm2[3] = {T(0), T(0), T(0)};  // h_{it-2} at the own node (fused, usteps = 2)
#pragma unroll
    for (int c = 0; c < 3; c++) {
        ai[c] = src[gi + c];
        if constexpr (FUSED) bi[c] = hold[gi + c];
#pragma unroll
        for (int k = 0; k < kShellPer; k++) {
            as_[k][c] = src[gs[k] + c];
            if constexpr (FUSED) bs_[k][c] = hold[gs[k] + c];
        }
    }
    if constexpr (FUSED) {
        if (ingrid && A.usteps != 0) {
#pragma unroll
            for (int c = 0; c < 3; c++) {
                ui[c] = (double)u[gi + c];
                hm2[c] = hnew[gi + c];  // before this apply overwrites it with h_it
            }
        }
    }

    // --- stage ---------------------------------------------------------------------
#pragma unroll
    for (int k = 0; k < kVoxPer; k++)
        if (tid + k * kSB < VN) ps[tid + k * kSB] = pvl[k];
    tps[tid] = (PermT)t;
    if (stop != 0) return;  // uniform over the block
    const int q = hpos_index(pin);
    int qs[kShellPer];
#pragma unroll
    for (int k = 0; k < kShellPer; k++) qs[k] = hpos_index(shp[k]);
    if constexpr (FUSED) {
        // this iteration's direction, h = z + gamma h_prev or h = z on a reset
        // (f90:1961-1996); the tile's own nodes also store it and, every other iteration,
        // take the two pending steps into u, u -= lambda h (f90:1914-1918), in order
        const bool dou = A.usteps != 0;
        auto dir = [&](T zc, T hc) -> T { return rst != 0 ? zc : (T)((double)zc + gam * (double)hc); };
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const T hv = dir(ai[c], bi[c]);
            hs[c * HN + q] = hv;
#pragma unroll
            for (int k = 0; k < kShellPer; k++)
                if (shv[k]) hs[c * HN + qs[k]] = dir(as_[k][c], bs_[k][c]);
            if (ingrid) {
                if (dou) {
                    const double u1 = ui[c] - lamp * (double)hm2[c];
                    u[gi + c] = (ustep_t<T>)(u1 - lam * (double)bi[c]);
                }
                hnew[gi + c] = hv;
            }
        }
    } else {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            hs[c * HN + q] = ai[c];
#pragma unroll
            for (int k = 0; k < kShellPer; k++)
                if (shv[k]) hs[c * HN + qs[k]] = as_[k][c];
        }
    }
    __syncthreads();

    // --- apply ---------------------------------------------------------------------
    const int tw = t >> 6, tl = t & 63;
    const int lx = wave_x(tw) + (tl & 3), ly = wave_y(tw) + ((tl >> 2) & 3), lz = wave_z(tw) + (tl >> 4);
    const bool active = ox + lx < nx && oy + ly < ny && oz + lz < nz;
    const int *pc = ps + (lz + 1) * (VX * VY) + (ly + 1) * VX + (lx + 1);
    int pv[8];
#pragma unroll
    for (int a = 0; a < 8; a++) pv[a] = pc[-(PZ(a) * (VX * VY) + PY(a) * VX + PX(a))];
    bool homog = true;
#pragma unroll
    for (int a = 1; a < 8; a++) homog &= pv[a] == pv[0];
    const bool hlane = active && homog;
    const T *hr = hs + (lz + 1) * (HX * HY) + (ly + 1) * HX + (lx + 1);

    T acc0 = T(0), acc1 = T(0), acc2 = T(0);
    if (__ballot(hlane)) {
        bool pend = hlane;
        while (true) {
            const unsigned long long pm = __ballot(pend);
            if (pm == 0ull) break;
            const int P = __builtin_amdgcn_readlane(pv[0], __builtin_ctzll(pm));
            ctab<T> *S = as_table(sten) + (int64_t)P * kStenHalf;
            if constexpr (sizeof(T) == 8) touch_lines<kStenHalf * 8>(sten + (int64_t)P * kStenHalf);
            T t0 = T(0), t1 = T(0), t2 = T(0);
#pragma unroll
            for (int k = 0; k < 13; k++) {
#pragma unroll
                for (int n = 0; n < 3; n++) {  // neighbour +delta
                    const T s0 = S[k * 9 + 0 * 3 + n], s1 = S[k * 9 + 1 * 3 + n], s2 = S[k * 9 + 2 * 3 + n];
                    const T hp = E3X_NB(hr, n, 14 + k);
                    t0 += s0 * hp; t1 += s1 * hp; t2 += s2 * hp;
                }
#pragma unroll
                for (int n = 0; n < 3; n++) {  // neighbour -delta: S[-d][j][n] = S[d][n][j]
                    const T s0 = S[k * 9 + n * 3 + 0], s1 = S[k * 9 + n * 3 + 1], s2 = S[k * 9 + n * 3 + 2];
                    const T hq = E3X_NB(hr, n, 12 - k);
                    t0 += s0 * hq; t1 += s1 * hq; t2 += s2 * hq;
                }
            }
            ctab<T> *Cc = S + 117;  // centre: 00 11 22 01 02 12
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
    // Mixed nodes, shared by the four waves: the tile's M mixed nodes sit at positions
    // [H, H + M) of the lane order.  Wave w applies the octant pair (w, w ^ 6) to each of
    // them, 64 at a time (lane l takes the node at position H + j0 + l), and leaves its
    // partial sum in LDS; after a barrier the lane holding the node adds the four partials
    // in wave order.  A tile's mixed work then runs on all four SIMDs at once instead of
    // keeping the block (and its LDS) resident while the one wave that holds the mixed
    // nodes goes through all eight octants.  Corner a ^ 6 is the one opposite a, and the
    // element matrix is symmetric under inversion through the voxel centre
    // (dk[P][a^6][i][b][j] = dk[P][a][i][b^6][j], the elasticity tensor being even), so
    // the second octant of a pair reads the first one's coefficient rows with its node
    // blocks taken in b ^ 6 order: the pair reads one slice, the second pass finds it in
    // the scalar cache.  For each octant the wave loops over the phases its lanes hold
    // there (ballot of the pending lanes, phase of the first one), so every lane is served
    // exactly once per octant and a node's sum runs in fixed order: results are bitwise
    // reproducible whatever the tile's other nodes hold.
    const unsigned hm = A.tcnt[(A.mstride != 0 ? (int64_t)r * ntile : 0) + bid];
    const int hcount = (int)(hm & 0xFFFFu), mcount = (int)(hm >> 16);
    if (mcount != 0) {  // uniform over the block
        static_assert(kSB == 256, "one octant pair per wave");
        // the wave's octant pair, uniform (so the coefficient loads stay scalar)
        const int a = __builtin_amdgcn_readfirstlane(tid >> 6), l = tid & 63;
#pragma unroll 1
        for (int j0 = 0; j0 < mcount; j0 += 64) {
            const int mi = j0 + l;
            const bool mine = mi < mcount;
            const int tm = tps[hcount + min(mi, mcount - 1)];
            const int mw = tm >> 6, ml = tm & 63;
            const int mx = wave_x(mw) + (ml & 3), my = wave_y(mw) + ((ml >> 2) & 3), mz = wave_z(mw) + (ml >> 4);
            const int *pcm = ps + (mz + 1) * (VX * VY) + (my + 1) * VX + (mx + 1);
            const T *hrm = hs + (mz + 1) * (HX * HY) + (my + 1) * HX + (mx + 1);
            T p0 = T(0), p1 = T(0), p2 = T(0);  // FP32: the pair's partial sum
#pragma unroll 1
            for (int half = 0; half < 2; half++) {
                const int oct = half ? (a ^ 6) : a;
                // voxel oct of the node spans nodes (PX(b)-PX(oct), ...), b = 0..7
                const T *hb = hrm - (PZ(oct) * (HX * HY) + PY(oct) * HX + PX(oct));
                T hv[24];
                if (half) {
#pragma unroll
                    for (int b = 0; b < 8; b++)
#pragma unroll
                        for (int n = 0; n < 3; n++)
                            hv[b * 3 + n] = hb[n * HN + PZ(b ^ 6) * (HX * HY) + PY(b ^ 6) * HX + PX(b ^ 6)];
                } else {
#pragma unroll
                    for (int b = 0; b < 8; b++)
#pragma unroll
                        for (int n = 0; n < 3; n++) hv[b * 3 + n] = hb[n * HN + PZ(b) * (HX * HY) + PY(b) * HX + PX(b)];
                }
                const int pa = pcm[-(PZ(oct) * (VX * VY) + PY(oct) * VX + PX(oct))];
                if constexpr (sizeof(T) == 8) {
                    // the slice of the octant's first phase (those of the others mostly
                    // follow from the octant before, or share its lines)
                    const int P0 = __builtin_amdgcn_readlane(pa, __builtin_ctzll(__ballot(mine)));
                    touch_lines<72 * 8>(dk + (int64_t)P0 * 576 + a * 72);
                }
                bool pend = mine;
                while (true) {
                    const unsigned long long pm = __ballot(pend);
                    if (pm == 0ull) break;
                    const int P = __builtin_amdgcn_readlane(pa, __builtin_ctzll(pm));
                    ctab<T> *D = as_table(dk) + (int64_t)P * 576 + a * 72;
                    T t0 = T(0), t1 = T(0), t2 = T(0);
#pragma unroll
                    for (int j = 0; j < 24; j++) {
                        const T d0 = D[j], d1 = D[24 + j], d2 = D[48 + j];
                        t0 += d0 * hv[j]; t1 += d1 * hv[j]; t2 += d2 * hv[j];
                    }
                    if (pend && pa == P) {
                        if constexpr (sizeof(T) == 8) {
                            // FP64: the pair's partial sum goes straight to LDS (six fewer
                            // live registers keep the body at 80, six waves per SIMD)
                            if (half) {
                                part[a][0][l] += t0; part[a][1][l] += t1; part[a][2][l] += t2;
                            } else {
                                part[a][0][l] = t0; part[a][1][l] = t1; part[a][2][l] = t2;
                            }
                        } else {
                            p0 += t0; p1 += t1; p2 += t2;
                        }
                        pend = false;
                    }
                }
            }
            if constexpr (sizeof(T) != 8) {
                part[a][0][l] = p0;
                part[a][1][l] = p1;
                part[a][2][l] = p2;
            }
            __syncthreads();
            const int own = tid - hcount - j0;
            if (own >= 0 && own < 64 && own + j0 < mcount) {
                acc0 = part[0][0][own]; acc1 = part[0][1][own]; acc2 = part[0][2][own];
#pragma unroll
                for (int w = 1; w < 4; w++) {
                    acc0 += part[w][0][own]; acc1 += part[w][1][own]; acc2 += part[w][2][own];
                }
            }
            if (j0 + 64 < mcount) __syncthreads();  // before the next round overwrites part
        }
    }

    // --- store ---------------------------------------------------------------------
    // The lane's node again from its lane-order entry (recomputed rather than kept live
    // through the FMA passes, which need the registers)
    int ts_ = t;
    asm volatile("" : "+v"(ts_));
    const int sw = ts_ >> 6, sl = ts_ & 63;
    const int sx = wave_x(sw) + (sl & 3), sy = wave_y(sw) + ((sl >> 2) & 3), sz = wave_z(sw) + (sl >> 4);
    const T *shr = hs + (sz + 1) * (HX * HY) + (sy + 1) * HX + (sx + 1);
    if constexpr (DIRECT) {
        // FP64 PCG apply: each lane stores its own node's Ah and leaves only the node's
        // h.Ah in LDS, at the node's x-fastest place; the last wave of the block sums those
        // as the tile epilogue below does (each lane over its four nodes in turn, then the
        // wave), so the partial is the same bits.  The block needs no Ah tile in LDS.
        __shared__ double prs[kSB];
        if (ox + sx < nx && oy + sy < ny && oz + sz < nz) {
            const int64_t m = o + 3 * (int64_t)((ox + sx) + nx * (oy + sy) + nxy * (oz + sz));
            y[m] = acc0;
            y[m + 1] = acc1;
            y[m + 2] = acc2;
        }
        double pr = 0.0;
        pr += (double)shr[0] * (double)acc0;
        pr += (double)shr[HN] * (double)acc1;
        pr += (double)shr[2 * HN] * (double)acc2;
        prs[sx + TX * (sy + TY * sz)] = pr;
        if (!last_wave<kSB>(&arrived)) return;
        const int l = tid & 63;
        double p0 = 0.0;
#pragma unroll
        for (int k = 0; k < kSB / 64; k++) {
            const int n = k * 64 + l;
            const int ex = n % TX, ey = (n / TX) % TY, ez = n / (TX * TY);
            if (ox + ex >= nx || oy + ey >= ny || oz + ez >= nz) continue;
            p0 += prs[n];
        }
        const double t0 = wave_sum(p0);
        if (l == 63) A.partials[(int64_t)r * ntile + bid] = t0;
        return;
    }
    // Ah into LDS at each lane's node in x-fastest tile order; the last wave of the block
    // to get here stores the tile row by row (the lanes hold their nodes in phase order)
    // and forms the partial sums, each lane over its four nodes in turn, then the wave.
    {
        const int nn = sx + TX * (sy + TY * sz);
        as[nn] = acc0; as[kSB + nn] = acc1; as[2 * kSB + nn] = acc2;
    }
    if (!last_wave<kSB>(&arrived)) return;
    int l = tid & 63;
    asm volatile("" : "+v"(l));  // keeps the tile addressing below from being hoisted above the FMA loops (spills)
    double p0 = 0.0, p1 = 0.0;  // h.Ah (or utot), gg (kEnergy)
#pragma unroll
    for (int k = 0; k < kSB / 64; k++) {
        const int n = k * 64 + l;
        const int ex = n % TX, ey = (n / TX) % TY, ez = n / (TX * TY);
        if (ox + ex >= nx || oy + ey >= ny || oz + ez >= nz) continue;
        const int64_t m = o + 3 * (int64_t)((ox + ex) + nx * (oy + ey) + nxy * (oz + ez));
        const T *h0 = hs + (ez + 1) * (HX * HY) + (ey + 1) * HX + (ex + 1);
        const T *an = as + n;
        if constexpr (MODE == kApply || MODE == kApplyFused) {
            double pr = 0.0;
#pragma unroll
            for (int j = 0; j < 3; j++) {
                const T aj = an[j * kSB];
                y[m + j] = aj;
                pr += (double)h0[j * HN] * (double)aj;
            }
            p0 += pr;
        } else if constexpr (MODE == kEnergy) {
            double e = 0.0, gsq = 0.0;
#pragma unroll
            for (int j = 0; j < 3; j++) {
                const int64_t qq = m + j;
                const double uq = (double)h0[j * HN], bq = bvec[qq], kx = (double)an[j * kSB];
                e = e + 0.5 * uq * kx + bq * uq;  // f90:1336
                const double gq = kx + bq;         // f90:1337
                y[qq] = (T)gq;
                gsq += gq * gq;                     // f90:452
            }
            p0 += e;
            p1 += gsq;
        } else if constexpr (MODE == kResid) {
#pragma unroll
            for (int j = 0; j < 3; j++) y32[m + j] = (float)((double)an[j * kSB] + bvec[m + j]);
        } else {
#pragma unroll
            for (int j = 0; j < 3; j++) y[m + j] = an[j * kSB];
        }
    }
    // partial slots [q * nrhs + r][tile]: h.Ah (or utot), then gg for kEnergy
    if constexpr (MODE == kApply || MODE == kApplyFused || MODE == kEnergy) {
        const double t0 = wave_sum(p0);
        double t1 = 0.0;
        if constexpr (MODE == kEnergy) t1 = wave_sum(p1);
        if (l == 63) {
            A.partials[(int64_t)r * ntile + bid] = t0;
            if constexpr (MODE == kEnergy) A.partials[(int64_t)(nrhs + r) * ntile + bid] = t1;
        }
    }
}
#undef E3X_NB

// At least WPE waves per SIMD: six for the FP64 body (80 VGPRs; forced into 64 VGPRs
// for eight waves it hoists fewer loads and ran 7 % slower), eight for the FP32 body.
// PhT: uint16_t phase ids, or uint32_t on a wide handle.
template <typename T, int MODE, int WPE, typename PhT>
__global__ __launch_bounds__(kSB) __attribute__((amdgpu_waves_per_eu(WPE, 8))) void k_stencil(StencilArgs<T> A)
{
    stencil_body<T, MODE, PhT>(A, static_cast<const PhT *>(A.pix), A.dk, A.sten, A.x, A.y, A.b, A.y32, A.z, A.hold,
                               A.hnew, A.u);
}

// Lane order of each stencil tile (Solver::perm), from the phases alone: one block per
// tile, thread t = node t of the tile's wave-cube numbering.  Key: homogeneous nodes
// first, by phase; then mixed nodes by (smallest phase, largest phase, the set of octants
// holding another phase than the smallest, octant-0 phase); nodes outside the grid last;
// ties by t.  A bitonic sort of the kSB keys in LDS.  The phases enter the key as their
// ranks among the tile's voxel-halo phases (a sorted copy in LDS, at most kSortN of
// them): the same order as the phases themselves, in 10 bits whatever their width.
constexpr int kSortN = 512;
static_assert(VN <= kSortN && kSortN % kSB == 0, "voxel halo must fit the rank table");

template <typename PhT>
__global__ __launch_bounds__(kSB) void k_tile_order(int nx, int ny, int nz, int ntx, int nty, const PhT *pix,
                                                  PermT *perm, unsigned *tcnt)
{
    __shared__ unsigned ps[VN];
    __shared__ unsigned srt[kSortN];
    __shared__ unsigned long long key[kSB];
    const int bid = blockIdx.x, t = threadIdx.x;
    const int tx = bid % ntx, ty = (bid / ntx) % nty, tz = bid / (ntx * nty);
    const int ox = tx * TX, oy = ty * TY, oz = tz * TZ, nxy = nx * ny;
    for (int q = t; q < kSortN; q += kSB) {
        unsigned v = 0xFFFFFFFFu;
        if (q < VN) {
            v = (unsigned)pix[pmod(ox + q % VX - 1, nx) + nx * pmod(oy + (q / VX) % VY - 1, ny) +
                              nxy * pmod(oz + q / (VX * VY) - 1, nz)];
            ps[q] = v;
        }
        srt[q] = v;
    }
    __syncthreads();
    for (int k = 2; k <= kSortN; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = t; i < kSortN; i += kSB) {
                const int p = i ^ j;
                if (p > i) {
                    const unsigned a = srt[i], b = srt[p];
                    if (((i & k) == 0) == (a > b)) {
                        srt[i] = b;
                        srt[p] = a;
                    }
                }
            }
            __syncthreads();
        }
    // rank of a phase: the position of its first copy in srt
    auto phase_rank = [&](unsigned v) -> unsigned {
        int lo = 0, hi = kSortN;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (srt[mid] < v) lo = mid + 1;
            else hi = mid;
        }
        return (unsigned)lo;
    };
    const int tw = t >> 6, tl = t & 63;
    const int lx = wave_x(tw) + (tl & 3), ly = wave_y(tw) + ((tl >> 2) & 3), lz = wave_z(tw) + (tl >> 4);
    const unsigned *pc = ps + (lz + 1) * (VX * VY) + (ly + 1) * VX + (lx + 1);
    unsigned pmin = 0xFFFFFFFFu, pmax = 0u, pv[8];
    const unsigned p0 = pc[0];
#pragma unroll
    for (int a = 0; a < 8; a++) {
        pv[a] = pc[-(PZ(a) * (VX * VY) + PY(a) * VX + PX(a))];
        pmin = min(pmin, pv[a]);
        pmax = max(pmax, pv[a]);
    }
    // which octants hold a phase other than the smallest: mixed nodes with the same phases
    // and the same pattern share every octant's phase, so the wave serves them in one pass
    // per octant
    unsigned sig = 0;
#pragma unroll
    for (int a = 0; a < 8; a++) sig |= (pv[a] != pmin ? 1u : 0u) << a;
    const bool active = ox + lx < nx && oy + ly < ny && oz + lz < nz;
    const unsigned long long cls = !active ? 2ull : (pmin == pmax ? 0ull : 1ull);
    key[t] = (cls << 62) | ((unsigned long long)phase_rank(pmin) << 50) |
             ((unsigned long long)phase_rank(pmax) << 40) | ((unsigned long long)sig << 32) |
             ((unsigned long long)phase_rank(p0) << 22) | (unsigned long long)t;
    // the tile's homogeneous and mixed node counts: per wave by ballot, then summed by
    // thread 0 after the barrier the key sort needs anyway
    __shared__ unsigned wcnt[kSB / 64][2];
    {
        const unsigned long long bh = __ballot(cls == 0ull), bm = __ballot(cls == 1ull);
        if ((t & 63) == 0) {
            wcnt[t >> 6][0] = (unsigned)__popcll(bh);
            wcnt[t >> 6][1] = (unsigned)__popcll(bm);
        }
    }
    __syncthreads();
    if (t == 0) {
        unsigned nhom = 0, nmix = 0;
        for (int w = 0; w < kSB / 64; w++) {
            nhom += wcnt[w][0];
            nmix += wcnt[w][1];
        }
        tcnt[bid] = nhom | (nmix << 16);
    }
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
    // Within each run of equal keys (same class and phases) any order serves: deal the run's
    // nodes out by their LDS bank residue (halo index mod 32), round by round, so that the
    // 16- and 32-lane groups that read LDS together hit distinct banks where the run allows
    // it.  Key: run start, rank of the node among the run's nodes of its residue, residue.
    static_assert(kSB <= 4096, "the run key holds 12-bit positions");
    __shared__ unsigned char resv[kSB];  // bank residue of the node at each sorted position
    const unsigned long long mine = key[t];
    const int node = (int)(mine & 0xFFFull);
    int res;
    {
        const int w = node >> 6, l = node & 63;
        const int x = wave_x(w) + (l & 3), y = wave_y(w) + ((l >> 2) & 3), z = wave_z(w) + (l >> 4);
        res = ((z + 1) * (HX * HY) + (y + 1) * HX + (x + 1)) & 31;
    }
    resv[t] = (unsigned char)res;
    // start of this lane's run: inclusive prefix maximum of the run heads
    __shared__ int sst[kSB];
    sst[t] = (t == 0 || (key[t - 1] >> 12) != (mine >> 12)) ? t : 0;
    __syncthreads();
    for (int off = 1; off < kSB; off <<= 1) {
        const int v = t >= off ? max(sst[t], sst[t - off]) : sst[t];
        __syncthreads();
        sst[t] = v;
        __syncthreads();
    }
    const int start = sst[t];
    int rank = 0;
    for (int i = start; i < t; i++) rank += resv[i] == res;
    __syncthreads();
    key[t] = ((unsigned long long)start << 40) | ((unsigned long long)rank << 24) | ((unsigned long long)res << 12) |
             (unsigned long long)node;
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
    perm[(int64_t)bid * kSB + t] = (PermT)(key[t] & 0xFFFull);
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
    A.d_ntile = fast_div((unsigned)stencil_blocks(s));
    A.d_ntx = fast_div((unsigned)A.ntx);
    A.d_nty = fast_div((unsigned)A.nty);
    A.wrap1 = s.nx >= 8 && s.ny >= 8 && s.nz >= 8;
    A.ns = s.ns; A.pix = s.pix; A.perm = s.perm; A.tcnt = s.tcnt; A.dk = dk; A.sten = sten; A.x = x; A.y = y;
    A.mstride = s.mstride();
    A.partials = s.partials; A.st = s.st;
    for (int r = 0; r < kMaxRhs; r++) A.rmap[r] = r;
    return A;
}

// One right-hand side per block (grid y = nrhs).  (Two FP32 right-hand sides per block,
// sharing the phase staging, classification and coefficient loads, ran 2-4 % slower at
// 200^3: the per-rhs FMA, LDS and staging work dominates, and the second halo costs
// occupancy.)
template <typename T, int MODE, int W>
void launch_stencil_w(const Solver &s, const StencilArgs<T> &A, dim3 grid)
{
    if (s.wide) hipLaunchKernelGGL((k_stencil<T, MODE, W, uint32_t>), grid, dim3(kSB), 0, s.stream, A);
    else hipLaunchKernelGGL((k_stencil<T, MODE, W, uint16_t>), grid, dim3(kSB), 0, s.stream, A);
}

// The PCG applies (kApply, kApplyFused) run at the occupancy the handle's tuning chose
// (tuned_waves: the waves per SIMD the compiler is held to, so their register budget);
// the other modes at the defaults, six waves for the FP64 body and eight for FP32.
template <typename T, int MODE>
void dispatch_stencil(const Solver &s, const StencilArgs<T> &A, int rows = -1)
{
    const dim3 grid((unsigned)stencil_blocks(s), (unsigned)(rows < 0 ? s.nrhs : rows));
    constexpr bool PCG = MODE == kApply || MODE == kApplyFused;
    const int w = PCG ? s.tuned_waves[sizeof(T) == 4 ? 1 : 0] : 0;
    if constexpr (sizeof(T) == 4) {
        if (PCG && w == 6) launch_stencil_w<T, MODE, 6>(s, A, grid);
        else launch_stencil_w<T, MODE, 8>(s, A, grid);
    } else {
        if (PCG && w == 5) launch_stencil_w<T, MODE, 5>(s, A, grid);
        else if (PCG && w == 8) launch_stencil_w<T, MODE, 8>(s, A, grid);
        else launch_stencil_w<T, MODE, 6>(s, A, grid);
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
// z and the previous one (h_alt / h by parity) while staging.  The grid has a row for
// each right-hand side not yet known (on the host) to have stopped; the blocks of one
// that stopped since still find that out from the device state.
template <typename T>
void apply_iter(Solver &s, const T *dk, const T *sten, T *hbuf[2], const T *z, T *ah, int it, int usteps,
                const int (&rmap)[kMaxRhs], int rows)
{
    StencilArgs<T> A = stencil_args<T>(s, dk, sten, it <= 1 ? hbuf[1] : nullptr, ah);
    for (int j = 0; j < kMaxRhs; j++) A.rmap[j] = rmap[j];
    if (it <= 1) {
        dispatch_stencil<T, kApply>(s, A, rows);
    } else {
        A.z = z; A.hold = hbuf[(it - 1) & 1]; A.hnew = hbuf[it & 1];
        if constexpr (sizeof(T) == 4) A.u = s.du32;
        else A.u = s.u;
        A.usteps = usteps;
        A.check_first = s.pcg_check_first ? 1 : 0;
        dispatch_stencil<T, kApplyFused>(s, A, rows);
    }
}

hipError_t launch_apply_hah(Solver &s, bool mixed, int it, int usteps)
{
    int rmap[kMaxRhs] = {}, rows = 0;
    for (int r = 0; r < s.nrhs; r++)
        if (!((s.pcg_known_stopped >> r) & 1u)) rmap[rows++] = r;
    if (rows == 0) return hipSuccess;  // every right-hand side has stopped: nothing to apply
    const bool timed = s.time_applies && it >= 2 && s.apply_n < kApplyTimed &&
                       (int)s.apply_ev.size() == 2 * kApplyTimed;
    if (timed) {
        hipError_t e = hipEventRecord(s.apply_ev[2 * s.apply_n], s.stream);
        if (e != hipSuccess) return e;
    }
    if (mixed) {
        float *hb[2] = {s.h32_alt, s.h32};
        apply_iter<float>(s, s.dk32, s.sten32, hb, s.z32, s.ah32, it, usteps, rmap, rows);
    } else {
        double *hb[2] = {s.h_alt, s.h};
        apply_iter<double>(s, s.dk, s.sten, hb, s.z, s.ah, it, usteps, rmap, rows);
    }
    if (timed) {
        hipError_t e = hipEventRecord(s.apply_ev[2 * s.apply_n + 1], s.stream);
        if (e != hipSuccess) return e;
        s.apply_it[s.apply_n++] = it;
    }
    return launch_hah_finish(s, (int)stencil_parts(s), it);
}

// The fused apply of iteration 2 with usteps u steps, as the PCG loop runs it (u and the
// state are left to the caller: with both lambdas and gamma 0 it leaves u unchanged and
// sets h = z).
hipError_t launch_apply_only(Solver &s, bool mixed, int usteps)
{
    int all[kMaxRhs];
    for (int r = 0; r < kMaxRhs; r++) all[r] = r;
    if (mixed) {
        float *hb[2] = {s.h32_alt, s.h32};
        apply_iter<float>(s, s.dk32, s.sten32, hb, s.z32, s.ah32, 2, usteps, all, s.nrhs);
    } else {
        double *hb[2] = {s.h_alt, s.h};
        apply_iter<double>(s, s.dk, s.sten, hb, s.z, s.ah, 2, usteps, all, s.nrhs);
    }
    return hipGetLastError();
}

hipError_t launch_tile_order(Solver &s, int member)
{
    const dim3 g((unsigned)stencil_blocks(s)), b(kSB);
    const int ntx = (s.nx + TX - 1) / TX, nty = (s.ny + TY - 1) / TY;
    const int64_t mb = (int64_t)member * s.mstride();                    // member's phase map
    PermT *perm = s.perm + (s.members ? (int64_t)member * stencil_blocks(s) * kSB : 0);
    unsigned *tcnt = s.tcnt + (s.members ? (int64_t)member * stencil_blocks(s) : 0);
    if (s.wide)
        hipLaunchKernelGGL(k_tile_order<uint32_t>, g, b, 0, s.stream, s.nx, s.ny, s.nz, ntx, nty,
                           (const uint32_t *)s.pix + mb, perm, tcnt);
    else
        hipLaunchKernelGGL(k_tile_order<uint16_t>, g, b, 0, s.stream, s.nx, s.ny, s.nz, ntx, nty,
                           (const uint16_t *)s.pix + mb, perm, tcnt);
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
