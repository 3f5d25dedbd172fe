// gfx950 PCG loop of dembx_OpenMP (src/elas3dxtal_pcg.f90:1356-2043).
//
// One iteration = four launches on the solver's stream, all gated on device-resident
// per-rhs stop state (so the host can enqueue iterations ahead and stop launching
// when it sees every right-hand side finished):
//   stencil  (it >= 2) h = z + gamma h_prev or z (f90:1961-1996) and, every other
//            iteration, the two pending u steps u -= lambda h (f90:1914-1918), then
//            Ah = A h and the partials of h.Ah         (e3x_stencil.hip)
//   fin_hah  hAh, |hAh| < 1e-14 break, lambda         (f90:1889-1907)
//   update   g -= lambda Ah, z = M^-1 g, partials of z.g (f90:1919-1946)
//   fin_upd  gg, rel_res, history, reset / gamma, stop test (f90:1949-2007)
// The u steps still pending and the direction of the last iteration are applied once
// after the loop (k_pcg_finalize).  The arithmetic of every value is the one of the separate steps.
// Storage type T of g, h, z, Ah and M^-1 is double, or float in the mixed form; u and
// every dot product stay double.
#include "e3x_device.hpp"

#include <atomic>
#include <cmath>

namespace e3x {
namespace {

// Cache hints of the streaming update (measured at 200^3): the FP32 form's read-once
// inputs (Ah, M^-1) are loaded non-temporally (-8 % for the kernel); in FP64 the same hint
// made it 14 % slower.  z is stored non-temporally in both forms: the next apply reads
// it through its halo lines later, and the hint also made that apply ~1 % faster.
template <typename T>
__device__ __forceinline__ T load_once(const T *p)
{
    if constexpr (sizeof(T) == 4) return __builtin_nontemporal_load(p);
    else return *p;
}

template <typename T>
struct PcgArgs {
    int64_t ns;
    int it, ldemb, kkk;
    int usteps;     // DO_U: 1, or 2 with the step of iteration it-1 (from hm1) first
    ustep_t<T> *u;  // u (FP64 form) or the FP32 u steps since the last replacement (mixed)
    T *g, *h, *z, *ah;
    const T *hm1;   // h_{it-1}
    const T *minv;
    const void *hphase;      // homogeneous phase of each node, all ones for the others
    int wide;                // hphase holds uint32 ids (else uint16)
    int64_t mstride;         // batch: nodes between the members' M^-1 and hphase arrays (0: shared)
    const T *minv_hom;       // [nphase][6]: M^-1 of a homogeneous node of each phase
    double *partials;
    PcgState *st;
    double *hist;
};

template <typename T>
__device__ __forceinline__ void precond_apply(const T *minv, int64_t ns, int64_t m, double g0, double g1,
                                              double g2, double &z0, double &z1, double &z2)
{
    const T *mi = minv + 6 * m;
    const double a00 = mi[0], a11 = mi[1], a22 = mi[2], a01 = mi[3], a02 = mi[4], a12 = mi[5];
    z0 = a00 * g0 + a01 * g1 + a02 * g2;
    z1 = a01 * g0 + a11 * g1 + a12 * g2;
    z2 = a02 * g0 + a12 * g1 + a22 * g2;
}

// z = M^-1 g, partials of z.g, and (kkk == 1) h = z  (f90:1527-1562)
template <typename T, int NR>
__global__ __launch_bounds__(kBlock) void k_pcg_init(PcgArgs<T> A)
{
    const int64_t ns = A.ns;
    double part[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) part[r] = 0.0;
    for (int64_t m = (int64_t)blockIdx.x * kBlock + threadIdx.x; m < ns; m += (int64_t)gridDim.x * kBlock) {
#pragma unroll
        for (int r = 0; r < NR; r++) {
            const int64_t o = ((int64_t)r * ns + m) * 3;
            T *g = A.g + o, *z = A.z + o, *h = A.h + o;
            const double g0 = g[0], g1 = g[1], g2 = g[2];
            double z0, z1, z2;
            precond_apply(A.minv + 6 * r * A.mstride, ns, m, g0, g1, g2, z0, z1, z2);
            const T t0 = (T)z0, t1 = (T)z1, t2 = (T)z2;
            z[0] = t0; z[1] = t1; z[2] = t2;
            if (A.kkk == 1) { h[0] = t0; h[1] = t1; h[2] = t2; }
            part[r] += (double)t0 * g0 + (double)t1 * g1 + (double)t2 * g2;
        }
    }
    write_partials<NR>(part, A.partials);
}

// loop-entry scalars (f90:1564-1567)
template <int NR>
__global__ __launch_bounds__(kFinBlock) void k_pcg_init_finish(const double *partials, int nparts, PcgState *st,
                                                            int ldemb)
{
    double tot[NR];
    sum_partials<NR>(partials, nparts, tot);
    if (threadIdx.x != 0) return;
    for (int r = 0; r < NR; r++) {
        PcgState &S = st[r];
        S.gg = tot[r];
        S.rel_res_prev = sqrt(S.gg / S.gg0);
        S.rel_res_min = S.rel_res_prev;
        S.rel_res = S.rel_res_prev;
        S.reset_count = 0;
        S.reset = 0;
        S.run_dir = 0;
        S.stop_iter = ldemb <= 0 ? -1 : 0;
        S.exit_flag = ldemb <= 0 ? 3 : 0;
    }
}

// hAh, the curvature guard and lambda (f90:1889-1907)
// hAh, the curvature guard and lambda (f90:1889-1907); block r serves right-hand side r
__global__ __launch_bounds__(kFinBlock) void k_hah_finish(const double *partials, int nparts, PcgState *st, int it)
{
    const int r = blockIdx.x;
    PcgState &S = st[r];
    if (S.stop_iter != 0) return;  // uniform over the block
    double tot[1];
    sum_partials<1>(partials + (int64_t)r * nparts, nparts, tot);
    if (threadIdx.x != 0) return;
    S.hah = tot[0];
    if (fabs(tot[0]) < 1e-14) {
        S.stop_iter = it;
        S.exit_flag = 2;  // E3X_PCG_HAH_BREAK
        S.run_dir = 0;
    } else {
        S.lambda_prev = S.lambda;
        S.lambda = S.gg / tot[0];
    }
}

// First stage of the h.Ah sum for large grids: block (g, r) adds its slice of right-hand
// side r's apply partials in a fixed order and writes out[r * gridDim.x + g] (a single
// finalize block alone keeps too few loads in flight for 31,250 partials).
__global__ __launch_bounds__(kBlock) void k_hah_stage(const double *partials, int nparts, const PcgState *st,
                                                    double *out)
{
    const int r = blockIdx.y, G = gridDim.x, g = blockIdx.x;
    if (st[r].stop_iter != 0) return;  // uniform over the block
    const int chunk = (nparts + G - 1) / G, lo = g * chunk, hi = min(nparts, lo + chunk);
    const double *p = partials + (int64_t)r * nparts;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = lo + (int)threadIdx.x; i < hi; i += 4 * kBlock) {
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int idx = i + u * kBlock;
            if (idx < hi) acc[u] += p[idx];
        }
    }
    double v[1] = {(acc[0] + acc[1]) + (acc[2] + acc[3])};
    block_sum<1, kBlock>(v);
    if (threadIdx.x == 0) out[(int64_t)r * G + g] = v[0];
}

// g -= lambda Ah (unless the true gradient replaces g this iteration); z = M^-1 g;
// partials of z.g (f90:1919-1946).  DO_U: u -= lambda h here (iterations whose true
// gradient needs the current u); otherwise the next iteration's apply moves u.  GSQ:
// also the partials of g.g, in slots [NR + r] (E3X_STOP_NIST_ABS).  MEM: the right-hand
// sides are the members of a batch, each with its own M^-1 (read per right-hand side).
template <typename T, int NR, bool UPDATE_G, bool DO_U, bool DO_Z, bool GSQ = false, bool MEM = false>
__device__ __forceinline__ void pcg_update_body(const PcgArgs<T> &A, ustep_t<T> *__restrict__ u, T *__restrict__ gv,
                                                const T *__restrict__ hv, T *__restrict__ zv,
                                                const T *__restrict__ ah, const T *__restrict__ minv)
{
    bool act[NR];
    double lam[NR], lamp[NR];
    bool any = false;
#pragma unroll
    for (int r = 0; r < NR; r++) {
        act[r] = A.st[r].stop_iter == 0;
        lam[r] = A.st[r].lambda;
        lamp[r] = A.st[r].lambda_prev;
        any |= act[r];
    }
    if (!any) return;
    const int64_t ns = A.ns;
    double part[GSQ ? 2 * NR : NR];
#pragma unroll
    for (int q = 0; q < (GSQ ? 2 * NR : NR); q++) part[q] = 0.0;
    for (int64_t m = (int64_t)blockIdx.x * kBlock + threadIdx.x; m < ns; m += (int64_t)gridDim.x * kBlock) {
        // the node's M^-1 block: from the per-phase table for a homogeneous node (the same
        // bits as its entry in minv), else from minv; once for all right-hand sides, or per
        // member of a batch
        double mi[6];
        auto load_mi = [&](int64_t n) {
            const int hp = hom_phase_at(A.hphase, A.wide, n);
            if (hp >= 0) {
#pragma unroll
                for (int q = 0; q < 6; q++) mi[q] = (double)A.minv_hom[6 * hp + q];
            } else {
#pragma unroll
                for (int q = 0; q < 6; q++) mi[q] = (double)load_once(&minv[6 * n + q]);
            }
        };
        if constexpr (DO_Z && !MEM) load_mi(m);
#pragma unroll
        for (int r = 0; r < NR; r++) {
            if (!act[r]) continue;
            const int64_t o = ((int64_t)r * ns + m) * 3;
            if constexpr (DO_Z && MEM) load_mi(r * A.mstride + m);
            if constexpr (DO_U) {
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    double uc = (double)u[o + c];
                    if (A.usteps == 2) uc = uc - lamp[r] * (double)A.hm1[o + c];
                    u[o + c] = (ustep_t<T>)(uc - lam[r] * (double)hv[o + c]);
                }
            }
            double g[3];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                g[c] = (double)gv[o + c];
                if constexpr (UPDATE_G) {
                    const T gn = (T)(g[c] - lam[r] * (double)load_once(&ah[o + c]));
                    gv[o + c] = gn;
                    g[c] = (double)gn;
                }
            }
            if constexpr (DO_Z) {
                // planes 00 11 22 01 02 12 (f90:1527-1533)
                const double z0 = mi[0] * g[0] + mi[3] * g[1] + mi[4] * g[2];
                const double z1 = mi[3] * g[0] + mi[1] * g[1] + mi[5] * g[2];
                const double z2 = mi[4] * g[0] + mi[5] * g[1] + mi[2] * g[2];
                const T t0 = (T)z0, t1 = (T)z1, t2 = (T)z2;
                __builtin_nontemporal_store(t0, &zv[o]);
                __builtin_nontemporal_store(t1, &zv[o + 1]);
                __builtin_nontemporal_store(t2, &zv[o + 2]);
                part[r] += (double)t0 * g[0] + (double)t1 * g[1] + (double)t2 * g[2];
                if constexpr (GSQ) part[NR + r] += g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
            }
        }
    }
    if constexpr (DO_Z) write_partials<GSQ ? 2 * NR : NR>(part, A.partials);
}

template <typename T, int NR, bool UPDATE_G, bool DO_U, bool DO_Z, bool GSQ = false, bool MEM = false>
__global__ __launch_bounds__(kBlock) void k_pcg_update(PcgArgs<T> A)
{
    pcg_update_body<T, NR, UPDATE_G, DO_U, DO_Z, GSQ, MEM>(A, A.u, A.g, A.h, A.z, A.ah, A.minv);
}

// gg, rel_res, history, divergence reset or gamma, stop test (f90:1938-2007); block r
// serves right-hand side r
__global__ __launch_bounds__(kFinBlock) void k_update_finish(const double *partials, int nparts, PcgState *st,
                                                          double *hist, int it, int nrhs)
{
    const int r = blockIdx.x;
    PcgState &S = st[r];
    if (S.stop_iter != 0) return;  // uniform over the block
    double tot[1], gsq[1] = {0.0};
    sum_partials<1>(partials + (int64_t)r * nparts, nparts, tot);
    if (S.stop_rule == 1) sum_partials<1>(partials + (int64_t)(nrhs + r) * nparts, nparts, gsq);  // uniform
    if (threadIdx.x != 0) return;
    if (S.stop_rule == 1) S.gsq = gsq[0];
    const double gglast = S.gg;
    S.gg = tot[0];
    const double rel = sqrt(S.gg / S.gg0);
    S.rel_res = rel;
    hist[(int64_t)r * S.ldemb + it - 1] = rel;
    if (S.no_reset == 0 && rel > 10.0 * S.rel_res_prev && it > 1) {
        S.reset = 1;
        S.rel_res_min = rel;
        S.reset_count += 1;
        S.gamma = 0.0;
    } else {
        S.reset = 0;
        S.gamma = S.gg / gglast;
        if (rel < S.rel_res_min) S.rel_res_min = rel;
    }
    S.rel_res_prev = rel;
    if (S.stop_rule == 1 ? gsq[0] < S.gsq_tol : rel < S.tol) {
        S.stop_iter = it;
        S.exit_flag = 1;  // E3X_PCG_CONVERGED
        S.run_dir = S.stop_first == 0;  // the variants exit before the direction update
    } else if (it >= S.ldemb) {
        S.stop_iter = it;
        S.exit_flag = 3;  // E3X_PCG_MAXITER
        S.run_dir = 1;
    }
}

// After the loop, per right-hand side that ended at iteration K, the u steps still
// pending (f90:1914-1918; the fused applies take them two at a time, u_done_after_apply):
// after convergence or the iteration limit the steps up to K, unless iteration K replaced
// the gradient (it then brought u up to K itself); after a curvature break (f90:1899-1902,
// no step K) those up to K-1.  Then the direction update that follows the stop test
// (f90:1961-1996, 2007) into h; where the loop ended before that update (run_dir 0: the
// sweep and plain-CG loops on convergence, A_2_0:1518, cg:1050) or broke, h_K itself.
// h_k sits in h (k odd) or h_alt (k even).
template <typename T, int NR>
__global__ __launch_bounds__(kBlock) void k_pcg_finalize(PcgArgs<T> A, T *h_alt, int replace_every)
{
    int K[NR], first[NR], last[NR];  // pending u steps first..last (none when first > last)
    int dir[NR];                     // 0 none, 1 direction update, 2 h_K into h
    bool rst[NR];
    double lamK[NR], lamK1[NR], gam[NR];  // lambdas of the steps K and K-1
    bool any = false;
#pragma unroll
    for (int r = 0; r < NR; r++) {
        const PcgState &S = A.st[r];
        const int k = S.stop_iter;
        K[r] = k;
        first[r] = 1;
        last[r] = 0;
        dir[r] = 0;
        lamK[r] = lamK1[r] = 0.0;
        if (k >= 1) {
            const bool conv = S.exit_flag == 1 || S.exit_flag == 3;
            first[r] = u_done_after_apply(k, replace_every) + 1;
            last[r] = conv ? ((replace_every > 0 && k % replace_every == 0) ? 0 : k) : k - 1;
            lamK[r] = conv ? S.lambda : 0.0;
            lamK1[r] = conv ? S.lambda_prev : S.lambda;  // a break leaves lambda at step K-1
            dir[r] = (conv && S.run_dir) ? 1 : 2;
        }
        rst[r] = S.reset != 0;
        gam[r] = S.gamma;
        any |= first[r] <= last[r] || dir[r] != 0;
    }
    if (!any) return;
    const int64_t ns = A.ns;
    for (int64_t m = (int64_t)blockIdx.x * kBlock + threadIdx.x; m < ns; m += (int64_t)gridDim.x * kBlock) {
#pragma unroll
        for (int r = 0; r < NR; r++) {
            if (first[r] > last[r] && dir[r] == 0) continue;
            const T *hk = (K[r] & 1) ? A.h : h_alt;    // h_K
            const T *hk1 = (K[r] & 1) ? h_alt : A.h;   // h_{K-1}
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const int64_t o = ((int64_t)r * ns + m) * 3 + c;
                const T hv = hk[o];
                if (first[r] <= last[r]) {
                    double uc = (double)A.u[o];
                    if (first[r] <= K[r] - 1) uc = uc - lamK1[r] * (double)hk1[o];  // step K-1
                    if (last[r] == K[r]) uc = uc - lamK[r] * (double)hv;            // step K
                    A.u[o] = (ustep_t<T>)uc;
                }
                if (dir[r] == 1) A.h[o] = rst[r] ? A.z[o] : (T)((double)A.z[o] + gam[r] * (double)hv);
                else if (dir[r] == 2 && hk != A.h) A.h[o] = hv;
            }
        }
    }
}

template <typename T>
PcgArgs<T> pcg_args(const Solver &s, int it)
{
    PcgArgs<T> A{};
    A.ns = s.ns; A.it = it;
    if constexpr (sizeof(T) == 4) A.u = s.du32;
    else A.u = s.u;
    A.partials = s.partials; A.st = s.st; A.hist = s.hist;
    A.hphase = s.hphase;
    A.wide = s.wide ? 1 : 0;
    A.mstride = s.mstride();
    if constexpr (sizeof(T) == 4) {
        A.g = s.g32; A.h = s.h32; A.z = s.z32; A.ah = s.ah32; A.minv = s.minv32; A.minv_hom = s.minv_hom32;
    } else {
        A.g = s.g; A.h = s.h; A.z = s.z; A.ah = s.ah; A.minv = s.minv; A.minv_hom = s.minv_hom;
    }
    return A;
}

template <typename T, int NR>
void init_nr(Solver &s, int kkk, int ldemb)
{
    const int g = (int)vector_grid(s);
    PcgArgs<T> A = pcg_args<T>(s, 0);
    A.kkk = kkk; A.ldemb = ldemb;
    hipLaunchKernelGGL((k_pcg_init<T, NR>), dim3(g), dim3(kBlock), 0, s.stream, A);
    hipLaunchKernelGGL(k_pcg_init_finish<NR>, dim3(1), dim3(kFinBlock), 0, s.stream, (const double *)s.partials, g,
                       s.st, ldemb);
}

// The streaming update's grid: at most one round of resident blocks.  With six
// right-hand sides the body takes 82 VGPRs (five waves per SIMD), and the 2048-block
// grid-stride grid would leave its last blocks a second, partly empty round.
template <typename T, int NR>
int update_grid(const Solver &s)
{
    // resident blocks on the device, per kernel instance; handles on several host threads
    // may race to fill it in, all with the same value
    static std::atomic<int> cached{0};
    int cap = cached.load(std::memory_order_relaxed);
    if (cap == 0) {
        int per_cu = 0, cu = 0, dev = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_pcg_update<T, NR, true, false, true>, kBlock, 0) !=
                hipSuccess ||
            per_cu <= 0)
            per_cu = kVecGridMax;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cu <= 0)
            cu = 1;
        cap = per_cu * cu;
        cached.store(cap, std::memory_order_relaxed);
    }
    const int64_t g = vector_grid(s);
    return (int)(g < cap ? g : cap);
}

// One launch of the streaming update in the handle's form: with the g.g partials (the NIST
// stop rule) or not, and the batch form (per-member M^-1) or not.
template <typename T, int NR, bool UG, bool DU, bool DZ>
void launch_update(const Solver &s, dim3 vg, const PcgArgs<T> &A)
{
    const dim3 blk(kBlock);
    const bool gsq = DZ && s.stop_rule == 1;
    if (s.members) {
        if (gsq) hipLaunchKernelGGL((k_pcg_update<T, NR, UG, DU, DZ, true, true>), vg, blk, 0, s.stream, A);
        else hipLaunchKernelGGL((k_pcg_update<T, NR, UG, DU, DZ, false, true>), vg, blk, 0, s.stream, A);
    } else {
        if (gsq) hipLaunchKernelGGL((k_pcg_update<T, NR, UG, DU, DZ, true>), vg, blk, 0, s.stream, A);
        else hipLaunchKernelGGL((k_pcg_update<T, NR, UG, DU, DZ>), vg, blk, 0, s.stream, A);
    }
}

template <typename T, int NR>
void iter_nr(Solver &s, int it, int replace_every)
{
    // a batch keeps the streaming grid of a single solve (the one-rhs grid is never capped),
    // so that every member's sums run over the same blocks as alone: the same bits
    const int g = s.members ? (int)vector_grid(s) : update_grid<T, NR>(s);
    const dim3 vg(g);
    const bool replace = replace_every > 0 && it % replace_every == 0;
    // 1. direction, the pending u steps, Ah = A h, the curvature (f90:1578-1907)
    (void)launch_apply_hah(s, sizeof(T) == 4, it, it >= 2 ? u_steps_of_apply(it, replace_every) : 0);
    PcgArgs<T> A = pcg_args<T>(s, it);
    // 2. updates, z and z.g (f90:1914-2007)
    if (!replace) {
        launch_update<T, NR, true, false, true>(s, vg, A);
    } else {
        // reliable update: u first (the pending step of iteration it-1, if any, then this
        // iteration's), then the true gradient g = K u + b in FP64, then z
        if constexpr (sizeof(T) == 4) {
            A.h = (it & 1) ? s.h32 : s.h32_alt;
            A.hm1 = (it & 1) ? s.h32_alt : s.h32;
        } else {
            A.h = (it & 1) ? s.h : s.h_alt;
            A.hm1 = (it & 1) ? s.h_alt : s.h;
        }
        A.usteps = it - u_done_after_apply(it, replace_every);
        launch_update<T, NR, false, true, false>(s, vg, A);
        if constexpr (sizeof(T) == 4) (void)launch_fold_du(s, true);  // u now has every step
        (void)launch_resid(s);
        launch_update<T, NR, false, false, true>(s, vg, A);
    }
    hipLaunchKernelGGL(k_update_finish, dim3(NR), dim3(kFinBlock), 0, s.stream, (const double *)s.partials, g, s.st,
                       s.hist, it, NR);
}

template <typename T, int NR>
void finalize_nr(Solver &s, int replace_every)
{
    PcgArgs<T> A = pcg_args<T>(s, 0);
    T *alt;
    if constexpr (sizeof(T) == 4) alt = s.h32_alt;
    else alt = s.h_alt;
    hipLaunchKernelGGL((k_pcg_finalize<T, NR>), dim3((unsigned)vector_grid(s)), dim3(kBlock), 0, s.stream, A, alt,
                       replace_every);
}

template <typename T>
void init_t(Solver &s, int kkk, int ldemb)
{
    switch (s.nrhs) {
    case 1: init_nr<T, 1>(s, kkk, ldemb); break;
    case 2: init_nr<T, 2>(s, kkk, ldemb); break;
    case 3: init_nr<T, 3>(s, kkk, ldemb); break;
    case 4: init_nr<T, 4>(s, kkk, ldemb); break;
    case 5: init_nr<T, 5>(s, kkk, ldemb); break;
    default: init_nr<T, 6>(s, kkk, ldemb); break;
    }
}

template <typename T>
void iter_t(Solver &s, int it, int replace_every)
{
    switch (s.nrhs) {
    case 1: iter_nr<T, 1>(s, it, replace_every); break;
    case 2: iter_nr<T, 2>(s, it, replace_every); break;
    case 3: iter_nr<T, 3>(s, it, replace_every); break;
    case 4: iter_nr<T, 4>(s, it, replace_every); break;
    case 5: iter_nr<T, 5>(s, it, replace_every); break;
    default: iter_nr<T, 6>(s, it, replace_every); break;
    }
}

template <typename T>
void finalize_t(Solver &s, int replace_every)
{
    switch (s.nrhs) {
    case 1: finalize_nr<T, 1>(s, replace_every); break;
    case 2: finalize_nr<T, 2>(s, replace_every); break;
    case 3: finalize_nr<T, 3>(s, replace_every); break;
    case 4: finalize_nr<T, 4>(s, replace_every); break;
    case 5: finalize_nr<T, 5>(s, replace_every); break;
    default: finalize_nr<T, 6>(s, replace_every); break;
    }
}

}  // namespace

hipError_t launch_hah_finish(Solver &s, int nparts, int it)
{
    const double *p = s.partials;
    constexpr int kStage = 64;
    if (nparts > 4096 && (int64_t)s.nrhs * (nparts + kStage) <= s.partials_cap) {
        // two stages: kStage partial sums per right-hand side behind the apply's partials
        double *stage = s.partials + (int64_t)s.nrhs * nparts;
        hipLaunchKernelGGL(k_hah_stage, dim3(kStage, (unsigned)s.nrhs), dim3(kBlock), 0, s.stream, p, nparts,
                           (const PcgState *)s.st, stage);
        p = stage;
        nparts = kStage;
    }
    hipLaunchKernelGGL(k_hah_finish, dim3((unsigned)s.nrhs), dim3(kFinBlock), 0, s.stream, p, nparts, s.st, it);
    return hipGetLastError();
}

hipError_t launch_pcg_init(Solver &s, bool mixed, int kkk, int ldemb, float tol, const double *gg0)
{
    // state scalars from the host (pinned staging): gg0, tol, ldemb
    PcgState *hs = s.st_host;
    for (int r = 0; r < s.nrhs; r++) {
        hs[r] = PcgState{};
        hs[r].gg0 = gg0[r];
        hs[r].tol = (double)tol;
        hs[r].ldemb = ldemb;
        hs[r].stop_rule = s.stop_rule;
        hs[r].no_reset = s.pcg_variant != 0;    // E3X_PCG_SWEEP, E3X_PCG_PLAIN_CG
        hs[r].stop_first = s.pcg_variant != 0;
        hs[r].gsq_tol = 1e-8 * (double)s.ns;  // NIST ELAS3D: gtest = 1e-8 * (number of voxels)
    }
    hipError_t e = hipMemcpyAsync(s.st, hs, sizeof(PcgState) * s.nrhs, hipMemcpyHostToDevice, s.stream);
    if (e != hipSuccess) return e;
    if (mixed) init_t<float>(s, kkk, ldemb);
    else init_t<double>(s, kkk, ldemb);
    return hipGetLastError();
}

hipError_t launch_pcg_iter(Solver &s, bool mixed, int it, int replace_every)
{
    if (mixed) iter_t<float>(s, it, replace_every);
    else iter_t<double>(s, it, 0);
    return hipGetLastError();
}

// Mixed form: u += du, du = 0 over the right-hand sides' rows (those still running, or all).
__global__ __launch_bounds__(kBlock) void k_fold_du(double *u, float *du, const PcgState *st, int64_t n3,
                                                  int running_only)
{
    const int r = blockIdx.y;
    if (running_only && st[r].stop_iter != 0) return;  // uniform over the block
    double *ur = u + (int64_t)r * n3;
    float *dr = du + (int64_t)r * n3;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n3; i += (int64_t)gridDim.x * kBlock) {
        ur[i] = ur[i] + (double)dr[i];
        dr[i] = 0.0f;
    }
}

hipError_t launch_fold_du(Solver &s, bool running_only)
{
    const int64_t n3 = 3 * s.ns;
    const int64_t g = std::min<int64_t>((n3 + kBlock - 1) / kBlock, 2048);
    hipLaunchKernelGGL(k_fold_du, dim3((unsigned)g, (unsigned)s.nrhs), dim3(kBlock), 0, s.stream, s.u, s.du32,
                       (const PcgState *)s.st, n3, running_only ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_pcg_finalize(Solver &s, bool mixed, int replace_every)
{
    if (mixed) {
        finalize_t<float>(s, replace_every);
        return launch_fold_du(s, false);  // u complete for every right-hand side
    }
    finalize_t<double>(s, 0);
    return hipGetLastError();
}

}  // namespace e3x
