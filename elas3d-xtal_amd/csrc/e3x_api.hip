// C ABI of the gfx950 elas3dxtal_pcg solver (include/e3x.h).
#include "e3x.h"
#include "e3x_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <tuple>
#include <string>
#include <vector>

using namespace e3x;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define E3X_HIP(call)                                                                           \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return fail(e_ == hipErrorOutOfMemory ? E3X_ENOMEM : E3X_EDEVICE,                   \
                        std::string(#call) + ": " + hipGetErrorString(e_));                     \
    } while (0)

// E3X_POISON_ALLOC=1 (read at each allocation; a debugging aid, off by default): every
// device allocation starts as 0xFF bytes (NaN in FP32 and FP64, the all-ones phase id),
// so that a kernel that reads a buffer before anything wrote it gives results that differ
// from a normal run's instead of depending on what the memory held before.
bool poison_alloc()
{
    const char *v = std::getenv("E3X_POISON_ALLOC");
    return v && v[0] == '1';
}

template <typename T>
hipError_t dalloc(Solver &s, T **p, int64_t n)
{
    *p = nullptr;
    if (n <= 0) return hipSuccess;
    hipError_t e = hipMalloc((void **)p, sizeof(T) * (size_t)n);
    if (e == hipSuccess) s.bytes += (int64_t)sizeof(T) * n;
    if (e == hipSuccess && poison_alloc()) {
        e = hipMemset(*p, 0xFF, sizeof(T) * (size_t)n);
        if (e == hipSuccess) e = hipDeviceSynchronize();
    }
    return e;
}

void release(Solver &s)
{
    (void)hipSetDevice(s.device);
    void *dev[] = {s.cmod, s.dk, s.sten, s.dk32, s.sten32, s.orient, s.pix, s.perm, s.tcnt, s.u, s.g, s.b, s.h,
                   s.h_alt, s.z, s.ah, s.minv, s.g32, s.h32, s.h32_alt, s.z32, s.ah32, s.minv32, s.du32, s.cconst,
                   s.partials, s.st, s.hist, s.results, s.post, s.hot, s.hphase, s.minv_hom, s.minv_hom32};
    for (void *p : dev)
        if (p) (void)hipFree(p);
    if (s.results_host) (void)hipHostFree(s.results_host);
    if (s.chk_host) (void)hipHostFree(s.chk_host);
    if (s.st_host) (void)hipHostFree(s.st_host);
    for (hipEvent_t &e : s.ev)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t &e : s.apply_ev)
        if (e) (void)hipEventDestroy(e);
    if (s.stream) (void)hipStreamDestroy(s.stream);
}

int check_handle(const e3x_solver *h)
{
    if (!h) return fail(E3X_EINVAL, "NULL solver handle");
    return E3X_OK;
}

// A handle whose PCG is in flight (e3x_dembx_begin .. e3x_dembx_end) takes no other call.
int check_idle(const e3x_solver *h)
{
    if (int rc = check_handle(h)) return rc;
    if (reinterpret_cast<const Solver *>(h)->run.on)
        return fail(E3X_ESTATE, "a PCG is in progress on this handle (e3x_dembx_end first)");
    return E3X_OK;
}

Solver &S(e3x_solver *h) { return *reinterpret_cast<Solver *>(h); }

int ensure_mixed_buffers(Solver &s)
{
    if (s.g32) return E3X_OK;
    const int64_t nv = (int64_t)s.nrhs * 3 * s.ns;
    E3X_HIP(dalloc(s, &s.g32, nv));
    E3X_HIP(dalloc(s, &s.h32, nv));
    E3X_HIP(dalloc(s, &s.h32_alt, nv));
    E3X_HIP(dalloc(s, &s.z32, nv));
    E3X_HIP(dalloc(s, &s.ah32, nv));
    E3X_HIP(dalloc(s, &s.minv32, 6 * s.ns * s.member_count()));
    E3X_HIP(dalloc(s, &s.du32, nv));
    // the buffers a fused apply reads before any PCG has written them (z and the
    // direction of iteration 1, e.g. under e3x_time_apply) start as zeros, not as stale
    // bits that lambda = 0 would turn into NaN
    E3X_HIP(hipMemsetAsync(s.z32, 0, sizeof(float) * (size_t)nv, s.stream));
    E3X_HIP(hipMemsetAsync(s.h32, 0, sizeof(float) * (size_t)nv, s.stream));
    E3X_HIP(hipMemsetAsync(s.h32_alt, 0, sizeof(float) * (size_t)nv, s.stream));
    E3X_HIP(hipMemsetAsync(s.du32, 0, sizeof(float) * (size_t)nv, s.stream));
    return E3X_OK;
}

int ensure_hist(Solver &s, int ldemb)
{
    const int64_t need = (int64_t)s.nrhs * std::max(ldemb, 1);
    if (need <= s.hist_cap) return E3X_OK;
    if (s.hist) {
        (void)hipFree(s.hist);
        s.bytes -= (int64_t)sizeof(double) * s.hist_cap;
        s.hist = nullptr;
    }
    E3X_HIP(dalloc(s, &s.hist, need));
    s.hist_cap = need;
    return E3X_OK;
}

// The launch tuning of the PCG apply, shared by the process's handles of one grid shape,
// right-hand-side layout and precision.
struct TuneKey {
    int nx, ny, nz, nrhs, members, wide, mixed;
    bool operator<(const TuneKey &o) const
    {
        return std::tie(nx, ny, nz, nrhs, members, wide, mixed) <
               std::tie(o.nx, o.ny, o.nz, o.nrhs, o.members, o.wide, o.mixed);
    }
};
std::mutex g_tune_mu;
std::map<TuneKey, std::pair<int, double>> g_tuned;

// Mean time of `reps` fused applies as the PCG runs them (the u steps alternating between
// none and two), the PCG state of every rhs set to run with lambda = gamma = 0 (u and z
// keep their values, h = z), then restored.  Overwrites h_alt / h32_alt, Ah and partials.
int time_fused_apply(Solver &s, bool mixed, int reps, double *avg_ms)
{
    PcgState saved[kMaxRhs];
    E3X_HIP(hipMemcpyAsync(s.st_host, s.st, sizeof(PcgState) * s.nrhs, hipMemcpyDeviceToHost, s.stream));
    E3X_HIP(hipStreamSynchronize(s.stream));
    for (int r = 0; r < s.nrhs; r++) {
        saved[r] = s.st_host[r];
        s.st_host[r].stop_iter = 0;
        s.st_host[r].lambda = 0.0;  // u stays as it is
        s.st_host[r].lambda_prev = 0.0;
        s.st_host[r].gamma = 0.0;   // h = z
        s.st_host[r].reset = 0;
    }
    E3X_HIP(hipMemcpyAsync(s.st, s.st_host, sizeof(PcgState) * s.nrhs, hipMemcpyHostToDevice, s.stream));
    E3X_HIP(launch_apply_only(s, mixed, 2));  // warm-up
    E3X_HIP(hipEventRecord(s.ev[0], s.stream));
    for (int i = 0; i < reps; i++) E3X_HIP(launch_apply_only(s, mixed, (i & 1) ? 2 : 0));
    E3X_HIP(hipEventRecord(s.ev[1], s.stream));
    E3X_HIP(hipEventSynchronize(s.ev[1]));
    float ms = 0;
    E3X_HIP(hipEventElapsedTime(&ms, s.ev[0], s.ev[1]));
    *avg_ms = ms / reps;
    for (int r = 0; r < s.nrhs; r++) s.st_host[r] = saved[r];
    E3X_HIP(hipMemcpyAsync(s.st, s.st_host, sizeof(PcgState) * s.nrhs, hipMemcpyHostToDevice, s.stream));
    E3X_HIP(hipStreamSynchronize(s.stream));
    return E3X_OK;
}

// Choose the waves per SIMD of the PCG apply in this precision: forced (e3x_set_tuning),
// already known for the shape, or measured now (each variant, a few launches on the
// handle's data; the fastest kept).  Every variant computes the same bits.
int tune_apply(Solver &s, bool mixed)
{
    const int k = mixed ? 1 : 0;
    if (s.tuned_waves[k] != 0) return E3X_OK;
    if (s.tune_force[k] != 0) {
        s.tuned_waves[k] = s.tune_force[k];
        return E3X_OK;
    }
    const TuneKey key{s.nx, s.ny, s.nz, s.nrhs, s.members, s.wide ? 1 : 0, k};
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tuned.find(key);
        if (it != g_tuned.end()) {
            s.tuned_waves[k] = it->second.first;
            s.tuned_apply_ms[k] = it->second.second;
            return E3X_OK;
        }
    }
    const int *cand = mixed ? kTuneWaves32 : kTuneWaves64;
    const int ncand = mixed ? 2 : 3;
    int best = 0;
    double best_ms = 0;
    for (int c = 0; c < ncand; c++) {
        s.tuned_waves[k] = cand[c];
        double ms = 0;
        if (int rc = time_fused_apply(s, mixed, 4, &ms)) {
            s.tuned_waves[k] = 0;
            return rc;
        }
        if (best == 0 || ms < best_ms) {
            best = cand[c];
            best_ms = ms;
        }
    }
    s.tuned_waves[k] = best;
    s.tuned_apply_ms[k] = best_ms;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned[key] = {best, best_ms};
    return E3X_OK;
}

}  // namespace

namespace e3x {

int solver_set_nphase(Solver &s, int nphase)
{
    if (s.run.on) return fail(E3X_ESTATE, "a PCG is in progress on this handle");
    if (nphase < 1 || nphase > s.nphase_cap)
        return fail(E3X_EINVAL, "nphase exceeds the phases this handle was created for");
    s.nphase = nphase;
    s.have_pix = s.have_femat = s.have_precond = s.have_precond32 = false;
    return E3X_OK;
}

// e3x_set_microstructure for phase ids already checked and converted (0-based, in the
// handle's width) with their histogram: no pass over the labels and no wait, for the
// request loop, whose preparation thread does that work.  ids and orientation must stay
// valid until the stream has run the copies.
int solver_set_microstructure_ids(Solver &s, const void *ids, bool wide_ids, const int64_t *counts,
                                  const double *orientation)
{
    if (s.run.on) return fail(E3X_ESTATE, "a PCG is in progress on this handle");
    if (s.members) return fail(E3X_ESTATE, "not for a batch handle (e3x_set_member)");
    if (!ids || !counts) return fail(E3X_EINVAL, "ids or counts is NULL");
    if (wide_ids != s.wide) return fail(E3X_EINVAL, "phase-id width does not match the handle");
    s.counts.assign(counts, counts + s.nphase);
    E3X_HIP(hipSetDevice(s.device));
    E3X_HIP(hipMemcpyAsync(s.pix, ids, (s.wide ? sizeof(uint32_t) : sizeof(uint16_t)) * (size_t)s.ns,
                           hipMemcpyHostToDevice, s.stream));
    if (orientation && s.nphase > 1)
        E3X_HIP(hipMemcpyAsync(s.orient, orientation, sizeof(double) * 3 * (size_t)(s.nphase - 1),
                               hipMemcpyHostToDevice, s.stream));
    E3X_HIP(launch_tile_order(s));
    s.have_pix = true;
    s.have_femat = s.have_precond = s.have_precond32 = false;
    return E3X_OK;
}

int report_error(int code, const char *msg)
{
    return fail(code, msg ? msg : "");
}

int solver_set_member_layout(Solver &s, const int *nphase)
{
    if (s.run.on) return fail(E3X_ESTATE, "a PCG is in progress on this handle");
    if (!s.members) return fail(E3X_ESTATE, "not a batch handle");
    int64_t tot = 0;
    for (int r = 0; r < s.members; r++) {
        if (nphase[r] < 1) return fail(E3X_EINVAL, "every member needs at least one phase");
        tot += nphase[r];
    }
    if (tot > s.nphase_cap) return fail(E3X_EINVAL, "the members' phases exceed the tables of this handle");
    int base = 0;
    for (int r = 0; r < s.members; r++) {
        s.phase_base[r] = base;
        s.member_nphase[r] = nphase[r];
        base += nphase[r];
    }
    s.nphase = base;
    s.member_set = 0;
    s.have_pix = s.have_femat = s.have_precond = s.have_precond32 = false;
    return E3X_OK;
}

}  // namespace e3x

extern "C" {

const char *e3x_last_error(void) { return g_err.c_str(); }

int e3x_version(void) { return 10000; }

int e3x_device_count(int *count)
{
    if (!count) return fail(E3X_EINVAL, "count is NULL");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    *count = e == hipSuccess ? n : 0;
    if (e != hipSuccess && e != hipErrorNoDevice) return fail(E3X_EDEVICE, hipGetErrorString(e));
    return E3X_OK;
}

int e3x_create(int nx, int ny, int nz, int nphase, int nrhs, int device, e3x_solver **out)
{
    if (!out) return fail(E3X_EINVAL, "out is NULL");
    *out = nullptr;
    Solver *s = nullptr;
    if (int rc = solver_create(nx, ny, nz, nphase, nrhs, 0, device, &s)) return rc;
    *out = reinterpret_cast<e3x_solver *>(s);
    return E3X_OK;
}

}  // extern "C"

namespace e3x {

int solver_create(int nx, int ny, int nz, int nphase, int nrhs, int members, int device, Solver **out)
{
    *out = nullptr;
    if (nx < 2 || ny < 2 || nz < 2) return fail(E3X_EINVAL, "grid extents must be >= 2");
    if ((int64_t)nx * ny * nz > (int64_t)INT32_MAX) return fail(E3X_EINVAL, "grid exceeds 2^31 nodes");
    if (nphase < 1 || nphase > kMaxPhases) return fail(E3X_EINVAL, "nphase must be in [1, 2^31 - 2]");
    if (nrhs < 1 || nrhs > kMaxRhs) return fail(E3X_EINVAL, "nrhs must be in [1, 6]");
    if (members != 0 && members != nrhs) return fail(E3X_EINVAL, "a batch has one right-hand side per member");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(E3X_EDEVICE, "no HIP device");
    if (device < 0 || device >= ndev) return fail(E3X_EINVAL, "device index out of range");
    Solver *s = new (std::nothrow) Solver();
    if (!s) return fail(E3X_ENOMEM, "host allocation failed");
    s->nx = nx; s->ny = ny; s->nz = nz; s->nphase = nphase; s->nphase_cap = nphase; s->nrhs = nrhs;
    s->members = members;
    s->wide = nphase > kNarrowPhases;
    s->device = device;
    s->ns = (int64_t)nx * ny * nz;
    auto bail = [&](int code) {
        release(*s);
        delete s;
        return code;
    };
#define CK(call)                                                                                \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return bail(fail(e_ == hipErrorOutOfMemory ? E3X_ENOMEM : E3X_EDEVICE,              \
                             std::string(#call) + ": " + hipGetErrorString(e_)));               \
    } while (0)
    CK(hipSetDevice(device));
    CK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    for (hipEvent_t &e : s->ev) CK(hipEventCreate(&e));
    const int64_t ns = s->ns, nv = (int64_t)nrhs * 3 * ns;
    CK(dalloc(*s, &s->cmod, (int64_t)nphase * 36));
    CK(dalloc(*s, &s->dk, (int64_t)nphase * 576));
    CK(dalloc(*s, &s->sten, (int64_t)nphase * kStenHalf));
    CK(dalloc(*s, &s->dk32, (int64_t)nphase * 576));
    CK(dalloc(*s, &s->sten32, (int64_t)nphase * kStenHalf));
    CK(dalloc(*s, &s->orient, (int64_t)3 * std::max(nphase - 1, 1)));
    const int64_t id_bytes = s->wide ? 4 : 2;  // phase-id width
    const int64_t nm = s->member_count();      // microstructures held
    CK(dalloc(*s, reinterpret_cast<uint8_t **>(&s->pix), id_bytes * ns * nm));
    CK(dalloc(*s, &s->perm, stencil_blocks(*s) * kSB * nm));
    CK(dalloc(*s, &s->tcnt, stencil_blocks(*s) * nm));
    CK(dalloc(*s, &s->u, nv));
    CK(dalloc(*s, &s->g, nv));
    CK(dalloc(*s, &s->b, nv));
    CK(dalloc(*s, &s->h, nv));
    CK(dalloc(*s, &s->h_alt, nv));
    CK(dalloc(*s, &s->z, nv));
    CK(dalloc(*s, &s->ah, nv));
    CK(dalloc(*s, &s->minv, 6 * ns * nm));
    CK(dalloc(*s, reinterpret_cast<uint8_t **>(&s->hphase), id_bytes * ns * nm));
    CK(dalloc(*s, &s->minv_hom, (int64_t)nphase * 6));
    CK(dalloc(*s, &s->minv_hom32, (int64_t)nphase * 6));
    CK(dalloc(*s, &s->cconst, kMaxRhs));
    // partials: [values][blocks] of the largest reduction (stress: 12 per rhs over the
    // streaming grid; energy: 2 per rhs over the stencil grid; rhs constant: 6)
    s->partials_cap = std::max<int64_t>(12 * (int64_t)nrhs * vector_grid(*s), 2 * (int64_t)nrhs * stencil_parts(*s));
    s->partials_cap = std::max<int64_t>(s->partials_cap, (int64_t)kMaxRhs * vector_grid(*s));
    CK(dalloc(*s, &s->partials, s->partials_cap));
    CK(dalloc(*s, &s->st, kMaxRhs));
    CK(dalloc(*s, &s->results, 12 * kMaxRhs));
    CK(hipHostMalloc((void **)&s->results_host, sizeof(double) * 12 * kMaxRhs, hipHostMallocDefault));
    CK(hipHostMalloc((void **)&s->chk_host, sizeof(PcgState) * 2 * kMaxRhs, hipHostMallocDefault));
    CK(hipHostMalloc((void **)&s->st_host, sizeof(PcgState) * 3 * kMaxRhs, hipHostMallocDefault));
    // fields read before anything writes them (h and z by a fused apply timed before any
    // PCG, e3x_time_apply) start as zeros.  The clears go on the handle's own stream and
    // are waited for: that stream is non-blocking, so a plain hipMemset (null stream, and
    // free to return before it completes) is ordered with none of the handle's later work
    // and could land after k_affine has written the initial guess into u.
    CK(hipMemsetAsync(s->h, 0, sizeof(double) * (size_t)nv, s->stream));
    CK(hipMemsetAsync(s->h_alt, 0, sizeof(double) * (size_t)nv, s->stream));
    CK(hipMemsetAsync(s->z, 0, sizeof(double) * (size_t)nv, s->stream));
    CK(hipMemsetAsync(s->u, 0, sizeof(double) * (size_t)nv, s->stream));
    CK(hipStreamSynchronize(s->stream));
#undef CK
    *out = s;
    return E3X_OK;
}

}  // namespace e3x

extern "C" {

int e3x_create_batch(int nx, int ny, int nz, int members, const int *nphase, int device, e3x_solver **out)
{
    if (!out) return fail(E3X_EINVAL, "out is NULL");
    *out = nullptr;
    if (members < 2 || members > kMaxRhs) return fail(E3X_EINVAL, "members must be in [2, 6]");
    if (!nphase) return fail(E3X_EINVAL, "nphase is NULL");
    int64_t tot = 0;
    for (int r = 0; r < members; r++) {
        if (nphase[r] < 1) return fail(E3X_EINVAL, "every member needs at least one phase");
        tot += nphase[r];
    }
    if (tot > kMaxPhases) return fail(E3X_EINVAL, "the members' phases exceed 2^31 - 2");
    Solver *s = nullptr;
    if (int rc = solver_create(nx, ny, nz, (int)tot, members, members, device, &s)) return rc;
    if (int rc = solver_set_member_layout(*s, nphase)) {
        release(*s);
        delete s;
        return rc;
    }
    *out = reinterpret_cast<e3x_solver *>(s);
    return E3X_OK;
}

int e3x_set_member(e3x_solver *h, int member, const int32_t *pix, const double *orientation, double c11,
                   double c12, double c44, const double *phasemod, int flag_m, const double *strain6)
{
    if (int rc = check_idle(h)) return rc;
    Solver &s = S(h);
    if (!s.members) return fail(E3X_ESTATE, "e3x_set_member needs a batch handle (e3x_create_batch)");
    if (member < 0 || member >= s.members) return fail(E3X_EINVAL, "member out of range");
    if (!pix || !strain6) return fail(E3X_EINVAL, "pix or strain6 is NULL");
    if (flag_m != 0 && flag_m != 1) return fail(E3X_EINVAL, "flag_m must be 0 or 1");
    const int np = s.member_nphase[member], base = s.phase_base[member];
    if (flag_m == 0 && np != 2) return fail(E3X_EINVAL, "flag_m = 0 needs nphase = 2");
    int32_t lo = INT32_MAX, hi = INT32_MIN;
    for (int64_t m = 0; m < s.ns; m++) {
        lo = std::min(lo, pix[m]);
        hi = std::max(hi, pix[m]);
    }
    if (lo < 1 || hi > np)  // f90:2986-2989, against the member's phases
        return fail(E3X_EPHASE, "phase label in pix out of bounds: min " + std::to_string(lo) + ", max " +
                                    std::to_string(hi) + ", nphase " + std::to_string(np));
    E3X_HIP(hipSetDevice(s.device));
    // the member's ids, shifted to its slice of the shared tables
    const int64_t n0 = (int64_t)member * s.ns;
    std::vector<uint16_t> p16;
    std::vector<uint32_t> p32;
    if (s.wide) {
        p32.resize((size_t)s.ns);
        for (int64_t m = 0; m < s.ns; m++) p32[(size_t)m] = (uint32_t)(pix[m] - 1 + base);
        E3X_HIP(hipMemcpyAsync(static_cast<uint32_t *>(s.pix) + n0, p32.data(), sizeof(uint32_t) * (size_t)s.ns,
                               hipMemcpyHostToDevice, s.stream));
    } else {
        p16.resize((size_t)s.ns);
        for (int64_t m = 0; m < s.ns; m++) p16[(size_t)m] = (uint16_t)(pix[m] - 1 + base);
        E3X_HIP(hipMemcpyAsync(static_cast<uint16_t *>(s.pix) + n0, p16.data(), sizeof(uint16_t) * (size_t)s.ns,
                               hipMemcpyHostToDevice, s.stream));
    }
    // orientations of the member's grains, at its slice (grains = phases less the pore)
    const double *ori = nullptr;
    if (orientation && np > 1) {
        E3X_HIP(hipMemcpyAsync(s.orient + 3 * (int64_t)base, orientation, sizeof(double) * 3 * (size_t)(np - 1),
                               hipMemcpyHostToDevice, s.stream));
        ori = s.orient + 3 * (int64_t)base;
    } else if (flag_m == 1 && np > 1) {
        return fail(E3X_EINVAL, "orientation is NULL");
    }
    E3X_HIP(launch_tile_order(s, member));
    // the ids staged in p16 / p32 (and the caller's orientations) must have reached the
    // device before they go out of scope: an asynchronous copy from pageable memory may
    // still be reading them when this call returns
    E3X_HIP(hipStreamSynchronize(s.stream));
    double pm[4] = {0, 0, 0, 0};
    if (phasemod) {
        pm[0] = phasemod[0];
        pm[1] = phasemod[1];
        if (flag_m == 0) {
            pm[2] = phasemod[2];
            pm[3] = phasemod[3];
        }
    }
    s.flag_m = flag_m;
    E3X_HIP(launch_femat_range(s, base, np, ori ? ori : s.orient, c11, c12, c44, pm));
    for (int q = 0; q < 6; q++) s.strain[6 * member + q] = strain6[q];
    s.member_set |= 1u << member;
    s.have_precond = s.have_precond32 = false;
    s.have_pix = s.have_femat = false;
    if (s.member_set == (1u << s.members) - 1u) {
        // every member is set: the jump right-hand sides and constants of all (f90:815-1100)
        E3X_HIP(launch_rhs(s));
        s.have_pix = s.have_femat = true;
    }
    E3X_HIP(hipStreamSynchronize(s.stream));
    return E3X_OK;
}

int e3x_destroy(e3x_solver *h)
{
    if (!h) return E3X_OK;
    Solver *s = reinterpret_cast<Solver *>(h);
    (void)hipSetDevice(s->device);
    (void)hipStreamSynchronize(s->stream);
    release(*s);
    delete s;
    return E3X_OK;
}

int e3x_memory_bytes(const e3x_solver *h, int64_t *bytes)
{
    if (int rc = check_handle(h)) return rc;
    if (!bytes) return fail(E3X_EINVAL, "bytes is NULL");
    *bytes = reinterpret_cast<const Solver *>(h)->bytes;
    return E3X_OK;
}

int e3x_set_microstructure(e3x_solver *h, const int32_t *pix, const double *orientation)
{
    if (int rc = check_idle(h)) return rc;
    Solver &s = S(h);
    if (s.members) return fail(E3X_ESTATE, "e3x_set_microstructure is not for a batch handle (e3x_set_member)");
    if (!pix) return fail(E3X_EINVAL, "pix is NULL");
    int32_t lo = INT32_MAX, hi = INT32_MIN;
    for (int64_t m = 0; m < s.ns; m++) {
        lo = std::min(lo, pix[m]);
        hi = std::max(hi, pix[m]);
    }
    // f90:2986-2989: labels must lie in [1, nphase]
    if (lo < 1 || hi > s.nphase)
        return fail(E3X_EPHASE, "phase label in pix out of bounds: min " + std::to_string(lo) + ", max " +
                                    std::to_string(hi) + ", nphase " + std::to_string(s.nphase));
    s.counts.assign((size_t)s.nphase, 0);
    for (int64_t m = 0; m < s.ns; m++) s.counts[(size_t)(pix[m] - 1)]++;
    E3X_HIP(hipSetDevice(s.device));
    // 0-based ids in the handle's width (16 bits up to kNarrowPhases phases, else 32)
    std::vector<uint16_t> p16;
    std::vector<uint32_t> p32;
    if (s.wide) {
        p32.resize((size_t)s.ns);
        for (int64_t m = 0; m < s.ns; m++) p32[(size_t)m] = (uint32_t)(pix[m] - 1);
        E3X_HIP(hipMemcpyAsync(s.pix, p32.data(), sizeof(uint32_t) * (size_t)s.ns, hipMemcpyHostToDevice, s.stream));
    } else {
        p16.resize((size_t)s.ns);
        for (int64_t m = 0; m < s.ns; m++) p16[(size_t)m] = (uint16_t)(pix[m] - 1);
        E3X_HIP(hipMemcpyAsync(s.pix, p16.data(), sizeof(uint16_t) * (size_t)s.ns, hipMemcpyHostToDevice, s.stream));
    }
    if (orientation && s.nphase > 1)
        E3X_HIP(hipMemcpyAsync(s.orient, orientation, sizeof(double) * 3 * (size_t)(s.nphase - 1),
                               hipMemcpyHostToDevice, s.stream));
    E3X_HIP(launch_tile_order(s));
    E3X_HIP(hipStreamSynchronize(s.stream));
    s.have_pix = true;
    s.have_femat = s.have_precond = s.have_precond32 = false;
    return E3X_OK;
}

int e3x_assig(e3x_solver *h, double *prob)
{
    if (int rc = check_idle(h)) return rc;
    Solver &s = S(h);
    if (s.members) return fail(E3X_ESTATE, "e3x_assig is not for a batch handle (e3x_set_member)");
    if (!prob) return fail(E3X_EINVAL, "prob is NULL");
    if (!s.have_pix) return fail(E3X_ESTATE, "e3x_assig before e3x_set_microstructure");
    // f90:2455-2465: counts of each label over ns (exact integers, counted on upload)
    for (int p = 0; p < s.nphase; p++) prob[p] = (double)s.counts[(size_t)p] / (double)s.ns;
    return E3X_OK;
}

int e3x_femat(e3x_solver *h, double c11, double c12, double c44, const double *phasemod, int flag_m,
              const double *strain6, double *C_out)
{
    if (int rc = check_idle(h)) return rc;
    Solver &s = S(h);
    if (s.members) return fail(E3X_ESTATE, "e3x_femat is not for a batch handle (e3x_set_member)");
    if (!s.have_pix) return fail(E3X_ESTATE, "e3x_femat before e3x_set_microstructure");
    if (flag_m != 0 && flag_m != 1) return fail(E3X_EINVAL, "flag_m must be 0 or 1");
    if (flag_m == 0 && s.nphase != 2) return fail(E3X_EINVAL, "flag_m = 0 needs nphase = 2");
    if (!strain6) return fail(E3X_EINVAL, "strain6 is NULL");
    double pm[4] = {0, 0, 0, 0};
    if (phasemod) {
        pm[0] = phasemod[0];
        pm[1] = phasemod[1];
        if (flag_m == 0) {
            pm[2] = phasemod[2];
            pm[3] = phasemod[3];
        }
    }
    s.flag_m = flag_m;
    for (int q = 0; q < 6 * kMaxRhs; q++) s.strain[q] = q < 6 * s.nrhs ? strain6[q] : 0.0;
    E3X_HIP(hipSetDevice(s.device));
    E3X_HIP(launch_femat(s, c11, c12, c44, pm));
    E3X_HIP(launch_rhs(s));
    if (C_out) {
        E3X_HIP(hipMemcpyAsync(s.results_host, s.cconst, sizeof(double) * s.nrhs, hipMemcpyDeviceToHost, s.stream));
        E3X_HIP(hipStreamSynchronize(s.stream));
        for (int r = 0; r < s.nrhs; r++) C_out[r] = s.results_host[r];
    } else {
        E3X_HIP(hipStreamSynchronize(s.stream));
    }
    s.have_femat = true;
    s.have_precond = s.have_precond32 = false;
    return E3X_OK;
}

int e3x_affine_guess(e3x_solver *h)
{
    if (int rc = check_idle(h)) return rc;
    Solver &s = S(h);
    if (!s.have_femat) return fail(E3X_ESTATE, "e3x_affine_guess before e3x_femat");
    E3X_HIP(hipSetDevice(s.device));
    E3X_HIP(launch_affine(s));
    E3X_HIP(hipStreamSynchronize(s.stream));
    return E3X_OK;
}

int e3x_energy(e3x_solver *h, double *utot, double *gg)
{
    if (int rc = check_idle(h)) return rc;
    Solver &s = S(h);
    if (!s.have_femat) return fail(E3X_ESTATE, "e3x_energy before e3x_femat");
    E3X_HIP(hipSetDevice(s.device));
    E3X_HIP(launch_energy(s));
    E3X_HIP(hipMemcpyAsync(s.results_host, s.results, sizeof(double) * 2 * s.nrhs, hipMemcpyDeviceToHost, s.stream));
    E3X_HIP(hipStreamSynchronize(s.stream));
    for (int r = 0; r < s.nrhs; r++) {
        if (utot) utot[r] = s.results_host[r];
        if (gg) gg[r] = s.results_host[s.nrhs + r];
    }
    return E3X_OK;
}

int e3x_set_stop_rule(e3x_solver *h, int rule)
{
    if (int rc = check_idle(h)) return rc;
    if (rule != E3X_STOP_REFERENCE && rule != E3X_STOP_NIST_ABS)
        return fail(E3X_EINVAL, "rule must be E3X_STOP_REFERENCE or E3X_STOP_NIST_ABS");
    S(h).stop_rule = rule;
    return E3X_OK;
}

int e3x_set_pcg_variant(e3x_solver *h, int variant)
{
    if (int rc = check_idle(h)) return rc;
    if (variant != E3X_PCG_REFERENCE && variant != E3X_PCG_SWEEP && variant != E3X_PCG_PLAIN_CG)
        return fail(E3X_EINVAL, "variant must be E3X_PCG_REFERENCE, E3X_PCG_SWEEP or E3X_PCG_PLAIN_CG");
    S(h).pcg_variant = variant;
    return E3X_OK;
}

int e3x_dembx_begin(e3x_solver *h, int precision, int replace_every, const double *gg0, float tol, int ldemb,
                    int kkk)
{
    if (int rc = check_handle(h)) return rc;
    Solver &s = S(h);
    if (!s.have_femat) return fail(E3X_ESTATE, "e3x_dembx before e3x_femat");
    if (s.run.on) return fail(E3X_ESTATE, "a PCG is already in progress on this handle");
    if (!gg0) return fail(E3X_EINVAL, "gg0 is NULL");
    if (precision != E3X_PREC_FP64 && precision != E3X_PREC_MIXED)
        return fail(E3X_EINVAL, "precision must be E3X_PREC_FP64 or E3X_PREC_MIXED");
    if (ldemb < 0) return fail(E3X_EINVAL, "ldemb must be >= 0");
    if (replace_every < 0) return fail(E3X_EINVAL, "replace_every must be >= 0");
    const bool mixed = precision == E3X_PREC_MIXED;
    E3X_HIP(hipSetDevice(s.device));
    if (mixed)
        if (int rc = ensure_mixed_buffers(s)) return rc;
    if (int rc = ensure_hist(s, ldemb)) return rc;
    const int64_t nv = (int64_t)s.nrhs * 3 * s.ns;
    if (int rc = tune_apply(s, mixed)) return rc;  // the apply's launch variant (once per shape)
    if (s.time_applies && s.apply_ev.empty()) {
        s.apply_ev.assign(2 * kApplyTimed, nullptr);
        s.apply_it.assign(kApplyTimed, 0);
        for (hipEvent_t &e : s.apply_ev) E3X_HIP(hipEventCreate(&e));
    }
    s.apply_n = 0;

    E3X_HIP(hipEventRecord(s.ev[0], s.stream));
    // block-Jacobi inverse, built once per femat (f90:1498-1520)
    const bool need32 = mixed && !s.have_precond32;
    const bool kind_changed = s.precond_identity != (s.pcg_variant == E3X_PCG_PLAIN_CG);
    if (!s.have_precond || kkk == 1 || need32 || kind_changed) {
        E3X_HIP(launch_precond(s, mixed || s.have_precond32));
        s.have_precond = true;
        if (mixed) s.have_precond32 = true;
    }
    E3X_HIP(hipEventRecord(s.ev[1], s.stream));
    if (mixed) {
        E3X_HIP(launch_to_f32(s.g, s.g32, nv, s.stream));
        if (kkk != 1) E3X_HIP(launch_to_f32(s.h, s.h32, nv, s.stream));
        // the u steps start from zero (a run that ended in an error may have left some)
        E3X_HIP(hipMemsetAsync(s.du32, 0, sizeof(float) * (size_t)nv, s.stream));
    }
    E3X_HIP(launch_pcg_init(s, mixed, kkk, ldemb, tol, gg0));
    Solver::PcgRun &R = s.run;
    R = Solver::PcgRun{};
    s.pcg_known_stopped = 0;
    s.pcg_check_first = false;
    R.on = true;
    R.mixed = mixed;
    R.replace_every = mixed ? replace_every : 0;
    R.ldemb = ldemb;
    R.chunk = s.ns >= (1 << 21) ? 4 : (s.ns >= (1 << 17) ? 16 : 32);
    return E3X_OK;
}

// The iteration at which every running right-hand side is expected to have stopped, from
// the slope of its log residual (the stop rule's quantity) between its last two snapshots;
// 0 when one of them cannot be predicted (no earlier snapshot, a residual that did not
// fall, as after a divergence reset, or a stop test that cannot be met).  It only sizes
// the chunks: the kernels take the stop decision on the device.
static int predict_stop(Solver &s, const PcgState *snap, int last)
{
    Solver::PcgRun &R = s.run;
    int pred = 0;
    bool all = true;
    for (int r = 0; r < s.nrhs; r++) {
        const PcgState &q = snap[r];
        if (q.stop_iter != 0) continue;
        const bool nist = q.stop_rule == 1;
        const double v = nist ? q.gsq : q.rel_res, target = nist ? q.gsq_tol : q.tol;
        int p = 0;
        if (v > 0.0 && target > 0.0 && std::isfinite(v)) {
            const double lv = std::log(v);
            if (R.seen_it[r] > 0 && last > R.seen_it[r]) {
                const double rate = (lv - R.seen_lv[r]) / (last - R.seen_it[r]);
                if (rate < 0.0) {
                    // iterations still to go, clamped before the conversion to int
                    const double need = std::min(std::max(std::ceil((std::log(target) - lv) / rate), 1.0), 1e6);
                    p = last + (int)need;
                }
            }
            R.seen_it[r] = last;
            R.seen_lv[r] = lv;
        }
        if (p == 0) all = false;
        pred = std::max(pred, p);
    }
    return all ? pred : 0;
}

// Every kernel of the loop reads the per-rhs stop state on the device, so the host only
// polls the state (double-buffered, one chunk behind) to stop launching.  A chunk ends one
// iteration past the predicted stop; past that, chunks are a quarter of the usual length.
int e3x_dembx_advance(e3x_solver *h, int block, int *done)
{
    if (int rc = check_handle(h)) return rc;
    Solver &s = S(h);
    Solver::PcgRun &R = s.run;
    if (!R.on) return fail(E3X_ESTATE, "e3x_dembx_advance without e3x_dembx_begin");
    if (!done) return fail(E3X_EINVAL, "done is NULL");
    *done = 0;
    E3X_HIP(hipSetDevice(s.device));
    while (!R.finishing) {
        // keep up to two chunks ahead of the last one examined
        while (!R.stop_seen && R.it <= R.ldemb && R.k - R.checked < 2) {
            int n = std::min(R.chunk, R.ldemb - R.it + 1);
            if (R.pred > 0) {
                const int left = R.pred + 1 - R.it + 1;  // through one iteration past the prediction
                n = std::min(n, left > 0 ? left : std::max(1, R.chunk / 4));
            }
            for (int t = 0; t < n; t++) {
                const int itt = R.it + t;
                s.pcg_check_first = R.pred > 0 && itt >= R.pred - 1;  // the prediction is good to ~2
                E3X_HIP(launch_pcg_iter(s, R.mixed, itt, R.mixed ? R.replace_every : 0));
            }
            s.pcg_check_first = false;
            R.it += n;
            // the state of every rhs into the pinned slot of this chunk
            R.chunk_last[R.k & 1] = R.it - 1;
            E3X_HIP(hipMemcpyAsync(s.chk_host + (R.k & 1) * kMaxRhs, s.st, sizeof(PcgState) * s.nrhs,
                                   hipMemcpyDeviceToHost, s.stream));
            E3X_HIP(hipEventRecord(s.ev[2 + (R.k & 1)], s.stream));
            R.k++;
        }
        if (R.checked < R.k) {  // examine the oldest chunk in flight
            hipEvent_t ev = s.ev[2 + (R.checked & 1)];
            if (block) {
                E3X_HIP(hipEventSynchronize(ev));
            } else {
                const hipError_t q = hipEventQuery(ev);
                if (q == hipErrorNotReady) return E3X_OK;
                E3X_HIP(q);
            }
            const PcgState *snap = s.chk_host + (R.checked & 1) * kMaxRhs;
            bool all = true;
            for (int r = 0; r < s.nrhs; r++) {
                const bool stopped = snap[r].stop_iter != 0;
                all &= stopped;
                if (stopped) s.pcg_known_stopped |= 1u << r;
            }
            R.pred = predict_stop(s, snap, R.chunk_last[R.checked & 1]);
            R.checked++;
            if (all) R.stop_seen = true;
            continue;
        }
        // every right-hand side stopped, or the iteration limit is enqueued and examined:
        // the last iteration's u step and direction update (deferred into the next apply)
        E3X_HIP(launch_pcg_finalize(s, R.mixed, R.replace_every));
        E3X_HIP(hipEventRecord(s.ev[2], s.stream));
        if (R.mixed) {
            const int64_t nv = (int64_t)s.nrhs * 3 * s.ns;
            E3X_HIP(launch_to_f64(s.g32, s.g, nv, s.stream));
            E3X_HIP(launch_to_f64(s.h32, s.h, nv, s.stream));
        }
        E3X_HIP(hipMemcpyAsync(s.st_host, s.st, sizeof(PcgState) * s.nrhs, hipMemcpyDeviceToHost, s.stream));
        E3X_HIP(hipEventRecord(s.ev[3], s.stream));
        R.finishing = true;
    }
    if (block) {
        E3X_HIP(hipEventSynchronize(s.ev[3]));
    } else {
        const hipError_t q = hipEventQuery(s.ev[3]);
        if (q == hipErrorNotReady) return E3X_OK;
        E3X_HIP(q);
    }
    *done = 1;
    return E3X_OK;
}

int e3x_dembx_end(e3x_solver *h, int *lstep, double *gg, double *history, int *exit_flag, int *reset_count)
{
    if (int rc = check_handle(h)) return rc;
    Solver &s = S(h);
    Solver::PcgRun &R = s.run;
    if (!R.on) return fail(E3X_ESTATE, "e3x_dembx_end without e3x_dembx_begin");
    int done = 0;
    if (int rc = e3x_dembx_advance(h, 1, &done)) {
        // the solve is abandoned: the handle takes calls again
        R = Solver::PcgRun{};
        s.pcg_known_stopped = 0;
        s.pcg_check_first = false;
        return rc;
    }
    R.on = false;
    const int ldemb = R.ldemb;
    float ms0 = 0, ms1 = 0;
    (void)hipEventElapsedTime(&ms0, s.ev[0], s.ev[1]);
    (void)hipEventElapsedTime(&ms1, s.ev[1], s.ev[2]);
    s.last_precond_ms = ms0;
    s.last_pcg_ms = ms1;
    // mean of the timed applies that did work: iterations up to the last stop (launches
    // past it find the stop state and leave at once)
    int last_stop = 0;
    for (int r = 0; r < s.nrhs; r++) last_stop = std::max(last_stop, s.st_host[r].stop_iter);
    double sum = 0;
    int n = 0;
    for (int k = 0; k < s.apply_n; k++) {
        if (s.apply_it[k] > last_stop) continue;
        float ms = 0;
        if (hipEventElapsedTime(&ms, s.apply_ev[2 * k], s.apply_ev[2 * k + 1]) == hipSuccess) {
            sum += ms;
            n++;
        }
    }
    s.last_apply_n = n;
    s.last_apply_ms = n > 0 ? sum / n : 0.0;
    s.apply_n = 0;
    for (int r = 0; r < s.nrhs; r++) {
        const PcgState &st = s.st_host[r];
        const int L = st.stop_iter > 0 ? st.stop_iter : 0;
        if (lstep) lstep[r] = L;
        if (gg) gg[r] = st.gg;
        if (exit_flag) exit_flag[r] = st.exit_flag;
        if (reset_count) reset_count[r] = st.reset_count;
        if (history) {
            if (L > 0)
                E3X_HIP(hipMemcpyAsync(history + (int64_t)r * ldemb, s.hist + (int64_t)r * ldemb,
                                       sizeof(double) * L, hipMemcpyDeviceToHost, s.stream));
            for (int q = L; q < ldemb; q++) history[(int64_t)r * ldemb + q] = NAN;
        }
    }
    if (history) E3X_HIP(hipStreamSynchronize(s.stream));
    return E3X_OK;
}

int e3x_dembx_ex(e3x_solver *h, int precision, int replace_every, const double *gg0, float tol, int ldemb,
                 int kkk, int *lstep, double *gg, double *history, int *exit_flag, int *reset_count)
{
    if (int rc = e3x_dembx_begin(h, precision, replace_every, gg0, tol, ldemb, kkk)) return rc;
    int done = 0;
    if (int rc = e3x_dembx_advance(h, 1, &done)) {
        S(h).run = Solver::PcgRun{};
        return rc;
    }
    return e3x_dembx_end(h, lstep, gg, history, exit_flag, reset_count);
}

int e3x_dembx(e3x_solver *h, const double *gg0, float tol, int ldemb, int kkk, int *lstep, double *gg,
              double *history, int *exit_flag, int *reset_count)
{
    return e3x_dembx_ex(h, E3X_PREC_FP64, 0, gg0, tol, ldemb, kkk, lstep, gg, history, exit_flag, reset_count);
}

int e3x_stress(e3x_solver *h, double *stress6, double *strain6)
{
    if (int rc = check_idle(h)) return rc;
    Solver &s = S(h);
    if (!s.have_femat) return fail(E3X_ESTATE, "e3x_stress before e3x_femat");
    E3X_HIP(hipSetDevice(s.device));
    E3X_HIP(launch_stress(s));
    E3X_HIP(hipMemcpyAsync(s.results_host, s.results, sizeof(double) * 12 * s.nrhs, hipMemcpyDeviceToHost, s.stream));
    E3X_HIP(hipStreamSynchronize(s.stream));
    for (int r = 0; r < s.nrhs; r++)
        for (int q = 0; q < 6; q++) {
            if (stress6) stress6[6 * r + q] = s.results_host[12 * r + q];
            if (strain6) strain6[6 * r + q] = s.results_host[12 * r + 6 + q];
        }
    return E3X_OK;
}

int e3x_stress_fullfield(e3x_solver *h, int rhs, double *stress, double *vm)
{
    if (int rc = check_idle(h)) return rc;
    Solver &s = S(h);
    if (!s.have_femat) return fail(E3X_ESTATE, "e3x_stress_fullfield before e3x_femat");
    if (rhs < 0 || rhs >= s.nrhs) return fail(E3X_EINVAL, "rhs out of range");
    E3X_HIP(hipSetDevice(s.device));
    if (!s.post) E3X_HIP(dalloc(s, &s.post, 7 * s.ns));  // staging: stress (6 ns) then vm (ns)
    double *dsf = stress ? s.post : nullptr, *dvm = vm ? s.post + 6 * s.ns : nullptr;
    E3X_HIP(launch_fullfield(s, rhs, dsf, dvm));
    if (stress)
        E3X_HIP(hipMemcpyAsync(stress, dsf, sizeof(double) * 6 * (size_t)s.ns, hipMemcpyDeviceToHost, s.stream));
    if (vm) E3X_HIP(hipMemcpyAsync(vm, dvm, sizeof(double) * (size_t)s.ns, hipMemcpyDeviceToHost, s.stream));
    E3X_HIP(hipStreamSynchronize(s.stream));
    return E3X_OK;
}

int e3x_hotspots(e3x_solver *h, int rhs, const double *vm, int nsolid, double sigma, double threshold_mpa,
                 int npeaks, int64_t *index, double *value, int64_t *found)
{
    if (int rc = check_idle(h)) return rc;
    Solver &s = S(h);
    if (!s.have_pix) return fail(E3X_ESTATE, "e3x_hotspots before e3x_set_microstructure");
    if (!vm && !s.have_femat) return fail(E3X_ESTATE, "e3x_hotspots without vm before e3x_femat");
    if (rhs < 0 || rhs >= s.nrhs) return fail(E3X_EINVAL, "rhs out of range");
    if (!(sigma > 0.0 && sigma <= 1.5)) return fail(E3X_EINVAL, "sigma must be in (0, 1.5]");
    if (npeaks < 0 || npeaks > kHotPeaksMax) return fail(E3X_EINVAL, "npeaks must be in [0, 64]");
    if (npeaks > 0 && (!index || !value)) return fail(E3X_EINVAL, "index or value is NULL");
    if (!std::isfinite(threshold_mpa)) return fail(E3X_EINVAL, "threshold_mpa must be finite");
    E3X_HIP(hipSetDevice(s.device));
    if (!s.post) E3X_HIP(dalloc(s, &s.post, 7 * s.ns));
    if (!s.hot) E3X_HIP(dalloc(s, &s.hot, 2 + 2 * (int64_t)kHotPeaksMax));
    double *dvm = s.post + 6 * s.ns;
    if (vm) E3X_HIP(hipMemcpyAsync(dvm, vm, sizeof(double) * (size_t)s.ns, hipMemcpyHostToDevice, s.stream));
    else E3X_HIP(launch_fullfield(s, rhs, nullptr, dvm));
    HotResult res;
    E3X_HIP(launch_hotspots(s, rhs, nsolid, sigma, (float)threshold_mpa, npeaks, res));
    for (size_t k = 0; k < res.peaks.size(); k++) {
        index[k] = res.peaks[k].m;
        value[k] = (double)res.peaks[k].v;
    }
    for (int k = (int)res.peaks.size(); k < npeaks; k++) {
        index[k] = -1;
        value[k] = 0.0;
    }
    if (found) *found = res.found;
    return E3X_OK;
}

int e3x_effective_stiffness(e3x_solver *h, double c11, double c12, double c44, const double *phasemod,
                            int flag_m, double magnitude, int precision, int replace_every, float tol,
                            int ldemb, double *C_eff, int *lstep, int *exit_flag)
{
    if (int rc = check_idle(h)) return rc;
    Solver &s = S(h);
    if (s.members) return fail(E3X_ESTATE, "e3x_effective_stiffness is not for a batch handle (e3x_set_member)");
    if (s.nrhs != 6) return fail(E3X_EINVAL, "e3x_effective_stiffness needs a handle with nrhs = 6");
    if (!C_eff) return fail(E3X_EINVAL, "C_eff is NULL");
    if (!(magnitude != 0.0) || !std::isfinite(magnitude)) return fail(E3X_EINVAL, "magnitude must be finite and nonzero");
    double e6[36] = {};
    for (int k = 0; k < 6; k++) e6[6 * k + k] = k < 3 ? magnitude : 0.5 * magnitude;
    if (int rc = e3x_femat(h, c11, c12, c44, phasemod, flag_m, e6, nullptr)) return rc;
    if (int rc = e3x_affine_guess(h)) return rc;
    double ut[6], gg0[6], gg[6];
    if (int rc = e3x_energy(h, ut, gg0)) return rc;
    int L[6], ex[6], resets[6];
    if (int rc = e3x_dembx_ex(h, precision, replace_every, gg0, tol, ldemb, 1, L, gg, nullptr, ex, resets))
        return rc;
    double sig[36], eps[36];
    if (int rc = e3x_stress(h, sig, eps)) return rc;
    // columns k of S and E are load case k: C E = S, solved as E^T C^T = S^T by
    // Gauss-Jordan elimination with partial pivoting
    double a[6][12];
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) {
            a[i][j] = eps[6 * i + j];       // E^T[i][j] = eps of case i, component j
            a[i][6 + j] = sig[6 * i + j];   // S^T[i][j] = sigma of case i, component j
        }
    for (int c = 0; c < 6; c++) {
        int p = c;
        for (int r = c + 1; r < 6; r++)
            if (std::fabs(a[r][c]) > std::fabs(a[p][c])) p = r;
        if (a[p][c] == 0.0) return fail(E3X_EINVAL, "averaged strains of the six load cases are singular");
        if (p != c)
            for (int j = 0; j < 12; j++) std::swap(a[p][j], a[c][j]);
        const double inv = 1.0 / a[c][c];
        for (int j = 0; j < 12; j++) a[c][j] *= inv;
        for (int r = 0; r < 6; r++) {
            if (r == c || a[r][c] == 0.0) continue;
            const double f = a[r][c];
            for (int j = 0; j < 12; j++) a[r][j] -= f * a[c][j];
        }
    }
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) C_eff[6 * i + j] = a[j][6 + i];  // C = (C^T)^T
    for (int k = 0; k < 6; k++) {
        if (lstep) lstep[k] = L[k];
        if (exit_flag) exit_flag[k] = ex[k];
    }
    return E3X_OK;
}

int e3x_get_field(e3x_solver *h, int which, int rhs, double *host)
{
    if (int rc = check_idle(h)) return rc;
    Solver &s = S(h);
    if (!host) return fail(E3X_EINVAL, "host is NULL");
    if (rhs < 0 || rhs >= s.nrhs) return fail(E3X_EINVAL, "rhs out of range");
    const double *src = which == E3X_FIELD_U ? s.u : which == E3X_FIELD_GB ? s.g : which == E3X_FIELD_B ? s.b
                        : which == E3X_FIELD_H ? s.h : nullptr;
    if (!src) return fail(E3X_EINVAL, "unknown field");
    E3X_HIP(hipSetDevice(s.device));
    // the device holds [ns][3] per right-hand side; the caller gets u(ns, 3).  Ah is free
    // outside a solve and serves as staging.
    E3X_HIP(launch_nodes_to_planes(src + (int64_t)rhs * 3 * s.ns, s.ah, s.ns, s.stream));
    E3X_HIP(hipMemcpyAsync(host, s.ah, sizeof(double) * 3 * (size_t)s.ns, hipMemcpyDeviceToHost, s.stream));
    E3X_HIP(hipStreamSynchronize(s.stream));
    return E3X_OK;
}

int e3x_set_field(e3x_solver *h, int which, int rhs, const double *host)
{
    if (int rc = check_idle(h)) return rc;
    Solver &s = S(h);
    if (!host) return fail(E3X_EINVAL, "host is NULL");
    if (rhs < 0 || rhs >= s.nrhs) return fail(E3X_EINVAL, "rhs out of range");
    double *dst = which == E3X_FIELD_U ? s.u : which == E3X_FIELD_GB ? s.g : which == E3X_FIELD_B ? s.b
                  : which == E3X_FIELD_H ? s.h : nullptr;
    if (!dst) return fail(E3X_EINVAL, "unknown field");
    E3X_HIP(hipSetDevice(s.device));
    E3X_HIP(hipMemcpyAsync(s.ah, host, sizeof(double) * 3 * (size_t)s.ns, hipMemcpyHostToDevice, s.stream));
    E3X_HIP(launch_planes_to_nodes(s.ah, dst + (int64_t)rhs * 3 * s.ns, s.ns, s.stream));
    E3X_HIP(hipStreamSynchronize(s.stream));
    return E3X_OK;
}

int e3x_get_tables(e3x_solver *h, double *cmod, double *dk)
{
    if (int rc = check_idle(h)) return rc;
    Solver &s = S(h);
    if (!s.have_femat) return fail(E3X_ESTATE, "e3x_get_tables before e3x_femat");
    E3X_HIP(hipSetDevice(s.device));
    if (cmod) E3X_HIP(hipMemcpyAsync(cmod, s.cmod, sizeof(double) * 36 * s.nphase, hipMemcpyDeviceToHost, s.stream));
    if (dk) E3X_HIP(hipMemcpyAsync(dk, s.dk, sizeof(double) * 576 * s.nphase, hipMemcpyDeviceToHost, s.stream));
    E3X_HIP(hipStreamSynchronize(s.stream));
    return E3X_OK;
}

int e3x_last_timing(e3x_solver *h, double *precond_ms, double *pcg_ms)
{
    if (int rc = check_handle(h)) return rc;
    Solver &s = S(h);
    if (precond_ms) *precond_ms = s.last_precond_ms;
    if (pcg_ms) *pcg_ms = s.last_pcg_ms;
    return E3X_OK;
}

int e3x_stress_dev(int nx, int ny, int nz, int nphase, int nrhs, const int32_t *pix, const double *cmod,
                   const double *strain6, const double *u, double *averages, void *stream)
{
    return stress_dev(nx, ny, nz, nphase, nrhs, pix, cmod, strain6, u, averages, (hipStream_t)stream);
}

int e3x_set_apply_timing(e3x_solver *h, int on)
{
    if (int rc = check_idle(h)) return rc;
    S(h).time_applies = on != 0;
    return E3X_OK;
}

int e3x_last_apply_timing(e3x_solver *h, int *launches, double *avg_ms)
{
    if (int rc = check_handle(h)) return rc;
    Solver &s = S(h);
    if (launches) *launches = s.last_apply_n;
    if (avg_ms) *avg_ms = s.last_apply_ms;
    return E3X_OK;
}

int e3x_set_tuning(e3x_solver *h, int precision, int waves)
{
    if (int rc = check_idle(h)) return rc;
    Solver &s = S(h);
    if (precision != E3X_PREC_FP64 && precision != E3X_PREC_MIXED)
        return fail(E3X_EINVAL, "precision must be E3X_PREC_FP64 or E3X_PREC_MIXED");
    const int k = precision == E3X_PREC_MIXED ? 1 : 0;
    bool ok = waves == 0;
    if (k == 0)
        for (int w : kTuneWaves64) ok |= w == waves;
    else
        for (int w : kTuneWaves32) ok |= w == waves;
    if (!ok) return fail(E3X_EINVAL, "waves must be 0 or a variant of the precision (FP64 5, 6, 8; FP32 6, 8)");
    s.tune_force[k] = waves;
    s.tuned_waves[k] = waves;  // 0: measured again at the next PCG
    s.tuned_apply_ms[k] = 0;
    return E3X_OK;
}

int e3x_get_tuning(e3x_solver *h, int precision, int *waves, double *apply_ms)
{
    if (int rc = check_handle(h)) return rc;
    Solver &s = S(h);
    const int k = precision == E3X_PREC_MIXED ? 1 : 0;
    if (waves) *waves = s.tuned_waves[k];
    if (apply_ms) *apply_ms = s.tuned_apply_ms[k];
    return E3X_OK;
}

int e3x_time_apply(e3x_solver *h, int precision, int reps, double *avg_ms)
{
    if (int rc = check_idle(h)) return rc;
    Solver &s = S(h);
    if (!s.have_femat) return fail(E3X_ESTATE, "e3x_time_apply before e3x_femat");
    if (reps < 1 || !avg_ms) return fail(E3X_EINVAL, "reps must be >= 1 and avg_ms non-NULL");
    const bool mixed = precision == E3X_PREC_MIXED;
    E3X_HIP(hipSetDevice(s.device));
    if (mixed)
        if (int rc = ensure_mixed_buffers(s)) return rc;
    if (int rc = tune_apply(s, mixed)) return rc;
    if (mixed) E3X_HIP(launch_to_f32(s.h, s.h32, (int64_t)s.nrhs * 3 * s.ns, s.stream));
    return time_fused_apply(s, mixed, reps, avg_ms);
}

}  // extern "C"
