// Internal declarations of the gfx950 elas3dxtal_pcg solver (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <type_traits>
#include <vector>

namespace e3x {

constexpr int kMaxRhs = 6;
// Phase ids are stored in 16 bits up to this many phases, in 32 bits above (the
// reference's anisotropy-sweep inputs have 81,488 grains); the all-ones id marks a node
// whose voxels hold several phases, so wide ids stay below 2^31 - 1.
constexpr int kNarrowPhases = 65535;
constexpr int kMaxPhases = 0x7FFFFFFE;
constexpr int kApplyTimed = 1024;  // fused applies timed per PCG (e3x_set_apply_timing)
constexpr int kBlock = 256;        // threads per block of every kernel
constexpr int kXcds = 8;           // MI355X: 8 XCDs, workgroups dispatched round-robin

// Stencil tile: 8 x 8 x 4 nodes per block of kSB threads; each 64-lane wave owns a
// 4x4x4 sub-cube.  (8 x 8 x 8 tiles, 512 threads, ran 8 % slower at 200^3.)
constexpr int TX = 8, TY = 8, TZ = 4;
constexpr int kSB = TX * TY * TZ;  // threads of a stencil block
using PermT = std::conditional_t<(kSB <= 256), uint8_t, uint16_t>;  // a lane's node in its tile
constexpr int HX = TX + 2, HY = TY + 2, HZ = TZ + 2, HN = HX * HY * HZ;  // node halo
constexpr int VX = TX + 1, VY = TY + 1, VZ = TZ + 1, VN = VX * VY * VZ;  // voxel halo
// Homogeneous 27-point stencil per phase, symmetric half: blocks of neighbours 14..26
// (9 each) then the centre block's upper triangle 00 11 22 01 02 12.
constexpr int kStenHalf = 13 * 9 + 6;

// What the u steps of a PCG in vectors of type T accumulate into: u itself (FP64), or in
// the mixed form the FP32 increment du32.
template <typename T>
using ustep_t = std::conditional_t<sizeof(T) == 4, float, double>;

// Per right-hand-side PCG scalars, resident on the device (the locals of
// dembx_OpenMP, f90:1440-1445).
struct PcgState {
    double gg, gg0, hah, lambda, gamma, rel_res, rel_res_prev, rel_res_min, tol;
    double lambda_prev;  // lambda of the iteration before (the u steps go two at a time)
    int stop_iter;    // 0 while running, else the iteration at which the loop ended
    int exit_flag;    // e3x_pcg_exit
    int reset;        // this iteration reset h = z (f90:1961-1981)
    int reset_count;
    int run_dir;      // the loop ended after this iteration's direction update (f90:2007 follows 1996)
    int ldemb;
    int stop_rule;    // e3x_stop_rule
    int no_reset;     // E3X_PCG_SWEEP, E3X_PCG_PLAIN_CG: no divergence reset
    int stop_first;   // E3X_PCG_SWEEP, E3X_PCG_PLAIN_CG: the stop test precedes the direction update
    int pad;
    double gsq_tol;   // E3X_STOP_NIST_ABS: stop when the sum of g^2 falls below this
    double gsq;       // the last iteration's sum of g^2 (E3X_STOP_NIST_ABS)
};

// The u steps u -= lambda_k h_k (f90:1914-1918) reach u two at a time: the fused apply of
// iteration it applies the two pending steps it-2 and it-1 when two are pending and none
// otherwise, which saves the u read and write of every other apply; a true-gradient
// replacement at iteration R (mixed form, every re iterations) brings u up to step R.
// Each step is still applied alone, in order, so u is the same value as with one step per
// iteration.  R0: the last replacement before iteration it (0 without any).
__host__ __device__ inline int u_last_replace(int it, int re) { return re > 0 ? ((it - 1) / re) * re : 0; }
// steps the fused apply of iteration it applies (0 or 2)
__host__ __device__ inline int u_steps_of_apply(int it, int re)
{
    const int d = it - u_last_replace(it, re);
    return (d >= 3 && (d & 1)) ? 2 : 0;
}
// the last step in u once the apply of iteration it has run
__host__ __device__ inline int u_done_after_apply(int it, int re)
{
    const int r0 = u_last_replace(it, re), d = it - r0;
    return r0 + 2 * ((d - 1) / 2);
}

// kApplyFused: the PCG apply of iterations >= 2, which first forms this iteration's
// direction h = z + gamma h_prev (h = z on a reset) and takes the previous step into u
// (u -= lambda h_prev), then applies K to the new h.
enum StencilMode { kApply = 0, kEnergy = 1, kResid = 2, kPlain = 3, kApplyFused = 4 };

struct Solver {
    int nx = 0, ny = 0, nz = 0, nphase = 0, nrhs = 1, device = 0;
    int nphase_cap = 0;  // phases the tables are allocated for (nphase <= nphase_cap)
    int64_t ns = 0;
    // A batch of separate problems of one grid size (e3x_create_batch; the serving loop's
    // co-batched requests): with members > 0, right-hand side r is a problem of its own,
    // with its own microstructure (pix, lane order, hphase and M^-1 at member offset r),
    // its phases at phase_base[r] .. phase_base[r] + member_nphase[r] - 1 of the shared
    // tables (pix holds those shifted ids), its materials and its load.  0: the
    // right-hand sides share one microstructure.
    int members = 0;
    int phase_base[kMaxRhs] = {}, member_nphase[kMaxRhs] = {};
    unsigned member_set = 0;  // members whose microstructure and tables are loaded (bit r)
    int64_t mstride() const { return members ? ns : 0; }  // nodes between members' per-node arrays
    int64_t member_count() const { return members ? members : 1; }
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {};
    int flag_m = 1;
    double strain[6 * kMaxRhs] = {};

    // phase tables
    double *cmod = nullptr, *dk = nullptr, *sten = nullptr;  // [p][36], [p][576], [p][kStenHalf]
    float *dk32 = nullptr, *sten32 = nullptr;
    double *orient = nullptr;                                  // [3*(nphase-1)]
    // 0-based phase per voxel: uint16 with at most kNarrowPhases phases, else uint32 (wide)
    void *pix = nullptr;
    bool wide = false;
    // per stencil tile, the order in which the block's lanes take its nodes (node t of
    // the tile's wave-cube numbering): homogeneous nodes by phase, then mixed nodes by
    // their phases, so that waves hold few distinct phases; built on upload
    PermT *perm = nullptr;                                     // [tiles][kSB]
    // per stencil tile, its homogeneous (low 16 bits) and mixed (high 16 bits) node counts:
    // the lane order holds the mixed nodes at positions [homogeneous, homogeneous + mixed)
    unsigned *tcnt = nullptr;                                  // [tiles]

    // fields, [nrhs][ns][3] (the three components of a node side by side)
    double *u = nullptr, *g = nullptr, *b = nullptr, *h = nullptr, *z = nullptr, *ah = nullptr;
    double *minv = nullptr;                                    // [ns][6]
    // homogeneous phase of each node (its 8 voxels share it), all ones for the others (the
    // width of pix), and the M^-1 block of a homogeneous node per phase ([nphase][6], FP64
    // and FP32)
    void *hphase = nullptr;
    double *minv_hom = nullptr;
    float *minv_hom32 = nullptr;
    float *g32 = nullptr, *h32 = nullptr, *z32 = nullptr, *ah32 = nullptr, *minv32 = nullptr;
    // mixed form: the u steps since the last true-gradient replacement, in FP32; folded
    // into u (FP64) before each replacement and after the loop (launch_fold_du)
    float *du32 = nullptr;
    // Second direction buffer: the direction of PCG iteration k lives in h (k odd) or
    // h_alt (k even), since the apply that forms it reads the previous one as halo.
    // h holds the final direction after e3x_dembx.
    double *h_alt = nullptr;
    float *h32_alt = nullptr;
    double *cconst = nullptr;                                  // C per rhs
    double *post = nullptr;                                    // full-field staging [7][ns]
    double *hot = nullptr;   // hotspot pass: two counters, then kHotPeaksMax entries

    // reductions and PCG state
    double *partials = nullptr;
    int64_t partials_cap = 0;
    PcgState *st = nullptr;
    PcgState *st_host = nullptr;     // pinned staging for the state upload / readback
    double *hist = nullptr;
    int64_t hist_cap = 0;
    double *results = nullptr;     // small device result slots
    double *results_host = nullptr;  // pinned mirror
    PcgState *chk_host = nullptr;    // pinned PCG state snapshots, two chunks' worth
    std::vector<int64_t> counts;     // phase histogram of pix (assig)

    // PCG in progress (e3x_dembx_begin / _advance / _end): chunks of iterations are
    // enqueued ahead; the state of every right-hand side after chunk k lands in
    // chk_host[(k & 1) * kMaxRhs + r], marked by ev[2 + (k & 1)]; at most two chunks are
    // unexamined at a time.  From the residuals of the chunks examined the host predicts
    // the stop iteration and ends the chunks there, so that few launches run past it.
    struct PcgRun {
        bool on = false, mixed = false, stop_seen = false, finishing = false;
        int replace_every = 0, ldemb = 0, chunk = 0, it = 1, k = 0, checked = 0;
        int chunk_last[2] = {0, 0};  // the last iteration of the chunk in slot k & 1
        int pred = 0;                // predicted stop iteration (0: none)
        int seen_it[kMaxRhs] = {};   // per rhs: iteration and log residual of the last snapshot
        double seen_lv[kMaxRhs] = {};
    } run;
    // right-hand sides the host has seen stop in the PCG in progress (bit r): the apply
    // launches no blocks for them
    uint32_t pcg_known_stopped = 0;
    // the iteration being enqueued lies past the predicted stop: its apply reads the stop
    // state before it loads its tile
    bool pcg_check_first = false;
    int stop_rule = 0;  // e3x_stop_rule of the next PCG
    int pcg_variant = 0;  // e3x_pcg_variant of the next PCG
    bool precond_identity = false;  // minv (minv32) hold the identity (E3X_PCG_PLAIN_CG)

    // e3x_set_apply_timing: HIP events around the fused applies (iterations >= 2) of the
    // following PCG runs, on the solver's stream; pair k brackets iteration apply_it[k]
    bool time_applies = false;
    std::vector<hipEvent_t> apply_ev;  // 2 * kApplyTimed, created on first use
    std::vector<int> apply_it;
    int apply_n = 0;                   // pairs recorded in the PCG in progress
    int last_apply_n = 0;
    double last_apply_ms = 0;          // mean over the launches of the last PCG
    // launch tuning of the PCG's stencil apply, [FP64, FP32]: waves per SIMD its register
    // budget is compiled for (0: not chosen yet; the tuner picks per grid shape, e3x_tuning)
    int tuned_waves[2] = {0, 0};
    int tune_force[2] = {0, 0};        // e3x_set_tuning: 0 measure, else the waves to use
    double tuned_apply_ms[2] = {0, 0};
    bool have_pix = false, have_femat = false, have_precond = false, have_precond32 = false;
    double last_precond_ms = 0, last_pcg_ms = 0;
    int64_t bytes = 0;
};

// ---- launchers --------------------------------------------------------------
// e3x_setup.hip
hipError_t launch_femat(Solver &s, double c11, double c12, double c44, const double *phasemod4);
hipError_t launch_rhs(Solver &s);                             // b and cconst
hipError_t launch_affine(Solver &s);
hipError_t launch_precond(Solver &s, bool want32);  // every member of a batch
// the tables of phases base .. base + nphase - 1 (a batch member's; 0, s.nphase otherwise)
hipError_t launch_femat_range(Solver &s, int base, int nphase, const double *orient, double c11, double c12,
                              double c44, const double *phasemod4);
// e3x_stencil.hip
hipError_t launch_energy(Solver &s);                          // results[0..2*nrhs)
hipError_t launch_apply_plain(Solver &s, const double *x, double *y);  // y = K x (FP64)
// PCG apply of iteration it (fused direction, and for it >= 2 the u steps usteps = 0 or
// 2 of u_steps_of_apply), then hAh and lambda
hipError_t launch_apply_hah(Solver &s, bool mixed, int it, int usteps);
hipError_t launch_apply_only(Solver &s, bool mixed, int usteps);  // fused apply with the current state
// the waves-per-SIMD variants of the PCG apply: FP64 {5, 6, 8}, FP32 {6, 8}
constexpr int kTuneWaves64[3] = {5, 6, 8}, kTuneWaves32[2] = {6, 8};
hipError_t launch_resid(Solver &s);                           // g32 = K u + b in FP64
hipError_t launch_tile_order(Solver &s, int member = 0);      // perm from pix (of a batch member)
// e3x_pcg.hip
hipError_t launch_hah_finish(Solver &s, int nparts, int it);
hipError_t launch_pcg_init(Solver &s, bool mixed, int kkk, int ldemb, float tol,
                           const double *gg0_host);
hipError_t launch_pcg_iter(Solver &s, bool mixed, int it, int replace_every);
// after the loop: the step still owed to u and the final direction, into h (h32)
hipError_t launch_pcg_finalize(Solver &s, bool mixed, int replace_every);
// mixed form: u += du32, du32 = 0 for the right-hand sides still running (running_only)
// or for all
hipError_t launch_fold_du(Solver &s, bool running_only);
// e3x_post.hip
hipError_t launch_stress(Solver &s);                          // results[0..12*nrhs)
hipError_t launch_fullfield(Solver &s, int rhs, double *stress_dev, double *vm_dev);
// e3x_stress_dev (caller-owned device arrays, scratch in e3x_post.hip); an e3x_status
int stress_dev(int nx, int ny, int nz, int nphase, int nrhs, const int32_t *pix, const double *cmod,
               const double *strain6, const double *u, double *averages, hipStream_t st);
hipError_t launch_to_f32(const double *src, float *dst, int64_t n, hipStream_t st);
hipError_t launch_to_f64(const float *src, double *dst, int64_t n, hipStream_t st);
// one right-hand side's field: component planes [3][ns] <-> node-interleaved [ns][3]
hipError_t launch_planes_to_nodes(const double *src, double *dst, int64_t ns, hipStream_t st);
hipError_t launch_nodes_to_planes(const double *src, double *dst, int64_t ns, hipStream_t st);

// e3x_hotspot.hip: regional maxima of the smoothed von Mises field in s.post's vm plane
struct HotEntry {
    float v;
    int pad;
    int64_t m;
};
constexpr int kHotPeaksMax = 64;
struct HotResult {
    int64_t found = 0;
    std::vector<HotEntry> peaks;
};
hipError_t launch_hotspots(Solver &s, int rhs, int nsolid, double sigma, float thr_mpa, int npeaks, HotResult &res);

// e3x_api.hip: reuse a handle for another microstructure with nphase <= nphase_cap
// phases (the serving loop's handle pool); the microstructure and tables must be
// loaded again.  Returns an e3x_status.
int solver_set_nphase(Solver &s, int nphase);
// e3x_set_microstructure from prepared 0-based phase ids (uint16, or uint32 on a wide
// handle) and their histogram; asynchronous (e3x_serve.hip)
int solver_set_microstructure_ids(Solver &s, const void *ids, bool wide_ids, const int64_t *counts,
                                  const double *orientation);
// e3x_api.hip: a handle for nrhs right-hand sides sharing one microstructure (members = 0)
// or for `members` separate problems (nrhs = members); an e3x_status
int solver_create(int nx, int ny, int nz, int nphase, int nrhs, int members, int device, Solver **out);
// e3x_api.hip: the phase counts of a batch's members (their slices of the tables, in
// order; the sum at most nphase_cap), before their e3x_set_member; an e3x_status
int solver_set_member_layout(Solver &s, const int *nphase);
// e3x_api.hip: set the message e3x_last_error() returns on this thread; returns code
int report_error(int code, const char *msg);

int64_t stencil_blocks(const Solver &s);
int64_t stencil_parts(const Solver &s);  // partials written per stencil launch (one per block)
inline int64_t vector_blocks(const Solver &s) { return (s.ns + kBlock - 1) / kBlock; }
// streaming kernels: grid-stride with at most kVecGridMax blocks (8 waves per SIMD)
constexpr int kVecGridMax = 2048;
inline int64_t vector_grid(const Solver &s) { return vector_blocks(s) < kVecGridMax ? vector_blocks(s) : kVecGridMax; }

}  // namespace e3x
