// Request loop over the gfx950 solver (include/e3x_serve.h): continuous batching of
// solves of different grid sizes on one GPU.
//
// Each active solve owns a solver handle, hence a HIP stream; its PCG advances in
// chunks (e3x_dembx_advance without waiting), so the kernels of all active solves are
// in flight together.  Queued requests of one small grid size (and the same precision
// and PCG controls) are admitted together as one batch handle (e3x_create_batch): their
// problems share every launch, which fills the GPU where one small grid alone leaves it
// idle.  Admission is by the device bytes a handle needs (first fit in arrival order
// under the budget and the slot limit); a finished handle stays in a pool for a later
// request (or batch) of the same grid size (freeing device memory synchronizes the
// device, so it is done only to make room).
#include "e3x.h"
#include "e3x_serve.h"
#include "e3x_internal.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <list>
#include <iterator>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

using namespace e3x;

namespace {

// Phases a serving handle is created for: any request with at most this many (grains
// plus the pore phase) can reuse it.
constexpr int kPoolPhases = 1024;
// Requests of at most this many nodes are prepared by e3x_server_submit itself; larger
// ones by the preparation thread, and progress admits them once they are ready.
constexpr int64_t kPrepInline = 64 * 64 * 64;

// A request's microstructure and what the server's preparation thread derives from it
// before admission: the label check (f90:2986-2989), the phase histogram and the 0-based
// phase ids in the width of the handle that will take it.  Shared with that thread.
struct Prep {
    std::vector<int32_t> pix;
    int nphase = 0;
    std::atomic<int> state{0};  // 0 pending, 1 ready, -1 a label out of [1, nphase]
    std::vector<uint16_t> id16;
    std::vector<uint32_t> id32;
    std::vector<int64_t> counts;
};

struct Request {
    int64_t id = 0;
    e3x_request rq{};
    std::shared_ptr<Prep> prep;
    std::vector<double> ori;
    double t_submit = 0;
};

// The preparation of one request: a single pass over its labels.
void prepare(Prep &p)
{
    const int64_t ns = (int64_t)p.pix.size();
    const bool wide = p.nphase > kNarrowPhases;
    p.counts.assign((size_t)p.nphase, 0);
    if (wide) p.id32.resize((size_t)ns);
    else p.id16.resize((size_t)ns);
    bool ok = true;
    for (int64_t m = 0; m < ns; m++) {
        const int32_t v = p.pix[(size_t)m];
        if (v < 1 || v > p.nphase) {
            ok = false;
            break;
        }
        p.counts[(size_t)(v - 1)]++;
        if (wide) p.id32[(size_t)m] = (uint32_t)(v - 1);
        else p.id16[(size_t)m] = (uint16_t)(v - 1);
    }
    if (!ok) {
        p.id16.clear();
        p.id32.clear();
    }
    p.state.store(ok ? 1 : -1, std::memory_order_release);
}

// The server's preparation thread: takes submitted requests in order, off the thread
// that drives the GPU (a 200^3 request is ~10 ms of host work).
struct Preparer {
    std::mutex mu;
    std::condition_variable cv;
    std::deque<std::shared_ptr<Prep>> todo;
    bool stop = false;
    std::thread th;

    Preparer()
    {
        th = std::thread([this] {
            for (;;) {
                std::shared_ptr<Prep> p;
                {
                    std::unique_lock<std::mutex> lk(mu);
                    cv.wait(lk, [this] { return stop || !todo.empty(); });
                    if (todo.empty()) return;  // stop, and nothing left
                    p = std::move(todo.front());
                    todo.pop_front();
                }
                prepare(*p);
            }
        });
    }
    ~Preparer()
    {
        {
            std::lock_guard<std::mutex> lk(mu);
            stop = true;
        }
        cv.notify_one();
        th.join();
    }
    void push(std::shared_ptr<Prep> p)
    {
        {
            std::lock_guard<std::mutex> lk(mu);
            todo.push_back(std::move(p));
        }
        cv.notify_one();
    }
};

struct Handle {
    e3x_solver *h = nullptr;
    int nx = 0, ny = 0, nz = 0, nphase_cap = 0;
    int members = 0;  // a batch handle's member count (0: a one-problem handle)
    int64_t bytes = 0;
};

// One solve on the GPU: a request on its own handle, or the members of a batch.
struct Active {
    std::vector<Request> reqs;
    Handle hd;
    bool small = false;  // a grid of at most batch_max_nodes nodes
    double t_admit = 0;
    double utot0[kMaxRhs] = {};
};

struct Server {
    int device = 0, max_active = 1;
    int64_t budget = 0, used = 0;
    std::chrono::steady_clock::time_point t0;
    int64_t next_id = 0;
    int batch_members = kMaxRhs;               // co-batching: at most this many per batch (1: off)
    int small_slots = 0;                       // slots of their own for grids of at most batch_max_nodes
    int64_t batch_max_nodes = 128 * 128 * 128; // ... of grids with at most this many nodes
    std::list<Request> queue;
    std::vector<Active> active;
    std::list<Handle> pool;  // idle handles, least recently used first
    std::deque<e3x_response> done;
    Preparer prep;           // label check and phase ids of submitted requests

    double now() const
    {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

Server &SV(e3x_server *p) { return *reinterpret_cast<Server *>(p); }
const Server &SV(const e3x_server *p) { return *reinterpret_cast<const Server *>(p); }

// The mixed form's FP32 buffers (ensure_mixed_buffers): g h h_alt z Ah du and M^-1.
int64_t mixed_bytes(int nx, int ny, int nz)
{
    const int64_t ns = (int64_t)nx * ny * nz;
    return 6 * 3 * ns * 4 + 6 * ns * 4;
}

// Device bytes of a one-rhs handle created for nphase_cap phases, with the mixed form's
// buffers and an iteration history of ldemb (the allocations of e3x_create,
// ensure_mixed_buffers and ensure_hist).
int64_t footprint(int nx, int ny, int nz, int nphase_cap, bool mixed, int ldemb)
{
    const int64_t ns = (int64_t)nx * ny * nz;
    const int64_t id = nphase_cap > kNarrowPhases ? 4 : 2;  // phase-id width
    int64_t b = ns * id;                                  // pix
    b += (int64_t)((nx + TX - 1) / TX) * ((ny + TY - 1) / TY) * ((nz + TZ - 1) / TZ) *
         (kSB * (int64_t)sizeof(PermT) + (int64_t)sizeof(unsigned));  // lane order and node counts
    b += 7 * 3 * ns * 8 + 6 * ns * 8 + ns * id;           // u g b h h_alt z Ah, M^-1, homogeneous phase
    b += (int64_t)nphase_cap * ((36 + 576 + kStenHalf + 3 + 6) * 8 + (576 + kStenHalf + 6) * 4);
    // partials (the larger of the streaming and the stencil reductions), state
    const int64_t tiles = (int64_t)((nx + TX - 1) / TX) * ((ny + TY - 1) / TY) * ((nz + TZ - 1) / TZ);
    b += std::max<int64_t>(12 * (int64_t)kVecGridMax, 2 * tiles) * 8 + (1 << 16);
    b += (int64_t)std::max(ldemb, 1) * 8;
    if (mixed) b += mixed_bytes(nx, ny, nz);
    return b;
}

int64_t handle_bytes(e3x_solver *h)
{
    int64_t b = 0;
    (void)e3x_memory_bytes(h, &b);
    return b;
}

// A handle's booking follows what it holds: its mixed buffers are allocated on its first
// mixed solve (the admission booked them beforehand, retire_handle settles the difference).
void retire_handle(Server &S, Handle hd)
{
    const int64_t now_bytes = handle_bytes(hd.h);
    S.used += now_bytes - hd.bytes;
    hd.bytes = now_bytes;
    S.pool.push_back(hd);
}

// A handle whose solve failed part way is not reused: its state is unknown.
void drop_handle(Server &S, Handle hd)
{
    S.used -= hd.bytes;
    (void)e3x_destroy(hd.h);
}

bool has_mixed_buffers(e3x_solver *h) { return reinterpret_cast<Solver *>(h)->g32 != nullptr; }

// Free idle handles, least recently used first, until `extra` more bytes fit.
void make_room(Server &S, int64_t extra)
{
    while (S.used + extra > S.budget && !S.pool.empty()) {
        S.used -= S.pool.front().bytes;
        (void)e3x_destroy(S.pool.front().h);
        S.pool.pop_front();
    }
}

// An idle handle of this grid size and enough phases, or one made room for and created.
// Every path books the bytes the request will hold before it is admitted: a pooled
// handle that lacks the mixed buffers a mixed request allocates is booked for them too.
int acquire_handle(Server &S, const e3x_request &rq, int members, int nphase_each, bool fields, int64_t need,
                   Handle *out, bool *fits)
{
    *fits = true;
    const bool mixed = rq.precision == E3X_PREC_MIXED;
    const int cap = std::max(kPoolPhases, nphase_each) * std::max(members, 1);
    for (auto it = S.pool.begin(); it != S.pool.end(); ++it)
        if (it->nx == rq.nx && it->ny == rq.ny && it->nz == rq.nz && it->members == members &&
            it->nphase_cap >= cap) {
            Handle hd = *it;
            const int64_t ns = (int64_t)rq.nx * rq.ny * rq.nz;
            const int64_t extra =
                ((mixed && !has_mixed_buffers(hd.h)) ? mixed_bytes(rq.nx, rq.ny, rq.nz) * std::max(members, 1) : 0) +
                ((fields && reinterpret_cast<Solver *>(hd.h)->post == nullptr) ? 7 * 8 * ns : 0);
            S.pool.erase(it);
            make_room(S, extra);
            if (S.used + extra > S.budget) {
                S.pool.push_front(hd);  // back among the idle handles; the request waits
                *fits = false;
                return E3X_OK;
            }
            S.used += extra;
            hd.bytes += extra;
            *out = hd;
            return E3X_OK;
        }
    // free idle handles of other sizes, least recently used first, until it fits
    make_room(S, need);
    if (S.used + need > S.budget) {
        *fits = false;
        return E3X_OK;
    }
    Handle hd;
    hd.nx = rq.nx; hd.ny = rq.ny; hd.nz = rq.nz;
    hd.nphase_cap = cap;
    hd.members = members;
    if (members == 0) {
        if (int rc = e3x_create(rq.nx, rq.ny, rq.nz, hd.nphase_cap, 1, S.device, &hd.h)) return rc;
    } else {
        Solver *sv = nullptr;
        if (int rc = solver_create(rq.nx, rq.ny, rq.nz, hd.nphase_cap, members, members, S.device, &sv)) return rc;
        hd.h = reinterpret_cast<e3x_solver *>(sv);
    }
    // booked at the admission estimate (which includes the mixed buffers and the history a
    // mixed request allocates later); retire_handle settles it to what the handle holds
    hd.bytes = std::max(handle_bytes(hd.h), need);
    S.used += hd.bytes;
    *out = hd;
    return E3X_OK;
}

e3x_response make_response(const Request &r, int status)
{
    e3x_response o;
    std::memset(&o, 0, sizeof(o));
    o.id = r.id;
    o.status = status;
    o.nx = r.rq.nx; o.ny = r.rq.ny; o.nz = r.rq.nz;
    o.t_submit = r.t_submit;
    return o;
}

// The preparation of a queued request: 0 pending, 1 ready, -1 a label out of [1, nphase]
// (f90:2986-2989; found before the request joins a solve, so that one bad request does
// not fail a batch).
int prep_state(const Request &r) { return r.prep->state.load(std::memory_order_acquire); }

// Requests that can share a batch: one grid, precision and set of PCG controls.
bool same_batch(const e3x_request &a, const e3x_request &b)
{
    return a.nx == b.nx && a.ny == b.ny && a.nz == b.nz && a.precision == b.precision &&
           a.replace_every == b.replace_every && a.tol == b.tol && a.ldemb == b.ldemb && a.flag_m == b.flag_m;
}

// The main program's sequence up to the PCG entry (f90:380-473) on the handle: for one
// request, or for every member of a batch.
int start_solve(Active &a)
{
    const e3x_request &q = a.reqs[0].rq;
    e3x_solver *h = a.hd.h;
    Solver &sv = *reinterpret_cast<Solver *>(h);
    const int n = (int)a.reqs.size();
    double gg0[kMaxRhs] = {};
    if (a.hd.members == 0) {
        if (int rc = solver_set_nphase(sv, q.nphase)) return rc;
        const Request &r = a.reqs[0];
        const Prep &p = *r.prep;
        const double *ori = r.ori.empty() ? nullptr : r.ori.data();
        // the prepared phase ids when their width is the handle's (a pooled handle made for
        // more than 65,535 phases takes 32-bit ids); else the labels themselves
        const bool wide_ids = !p.id32.empty();
        if (wide_ids == sv.wide && (wide_ids || !p.id16.empty())) {
            const void *ids = wide_ids ? (const void *)p.id32.data() : (const void *)p.id16.data();
            if (int rc = solver_set_microstructure_ids(sv, ids, wide_ids, p.counts.data(), ori)) return rc;
        } else if (int rc = e3x_set_microstructure(h, p.pix.data(), ori)) {
            return rc;
        }
        double cval = 0;
        if (int rc = e3x_femat(h, q.c11, q.c12, q.c44, q.phasemod, q.flag_m, q.strain6, &cval)) return rc;
    } else {
        int np[kMaxRhs];
        for (int k = 0; k < n; k++) np[k] = a.reqs[(size_t)k].rq.nphase;
        if (int rc = solver_set_member_layout(sv, np)) return rc;
        for (int k = 0; k < n; k++) {
            const Request &r = a.reqs[(size_t)k];
            const e3x_request &rk = r.rq;
            if (int rc = e3x_set_member(h, k, r.prep->pix.data(), r.ori.empty() ? nullptr : r.ori.data(), rk.c11, rk.c12,
                                        rk.c44, rk.phasemod, rk.flag_m, rk.strain6))
                return rc;
        }
    }
    if (int rc = e3x_affine_guess(h)) return rc;
    if (int rc = e3x_energy(h, a.utot0, gg0)) return rc;
    return e3x_dembx_begin(h, q.precision, q.replace_every, gg0, q.tol, q.ldemb, 1);
}

// After the PCG (f90:479-519): energy and averages, one response per request.
int finish_solve(Active &a, std::vector<e3x_response> &out)
{
    e3x_solver *h = a.hd.h;
    const size_t n = a.reqs.size();
    int lstep[kMaxRhs], ex[kMaxRhs], rs[kMaxRhs];
    double gg[kMaxRhs], ut[kMaxRhs], g2[kMaxRhs], st[6 * kMaxRhs], sn[6 * kMaxRhs];
    if (int rc = e3x_dembx_end(h, lstep, gg, nullptr, ex, rs)) return rc;
    if (int rc = e3x_energy(h, ut, g2)) return rc;
    if (int rc = e3x_stress(h, st, sn)) return rc;
    for (size_t k = 0; k < n; k++) {
        const e3x_request &q = a.reqs[k].rq;
        if (q.stress_field || q.vm)  // the optional full-field response (f90:2639-2833)
            if (int rc = e3x_stress_fullfield(h, (int)k, q.stress_field, q.vm)) return rc;
    }
    for (size_t k = 0; k < n; k++) {
        e3x_response &o = out[k];
        o.lstep = lstep[k];
        o.gg = gg[k];
        o.exit_flag = ex[k];
        o.reset_count = rs[k];
        o.utot = ut[k];
        o.utot0 = a.utot0[k];
        for (int q = 0; q < 6; q++) {
            o.stress6[q] = st[6 * k + q];
            o.strain6[q] = sn[6 * k + q];
        }
    }
    return E3X_OK;
}

void respond(Server &S, const Active &a, std::vector<e3x_response> &out, int status)
{
    for (e3x_response &o : out) {
        if (status != E3X_OK) o.status = status;
        o.t_admit = a.t_admit;
        o.t_done = S.now();
        S.done.push_back(o);
    }
}

int progress(Server &S, bool *changed)
{
    *changed = false;
    // retire the solves whose PCG has finished on the device
    for (size_t i = 0; i < S.active.size();) {
        Active &a = S.active[i];
        int fin = 0;
        int rc = e3x_dembx_advance(a.hd.h, 0, &fin);
        if (rc == E3X_OK && !fin) {
            ++i;
            continue;
        }
        std::vector<e3x_response> out;
        for (const Request &r : a.reqs) out.push_back(make_response(r, E3X_OK));
        if (rc == E3X_OK) rc = finish_solve(a, out);
        else (void)e3x_dembx_end(a.hd.h, nullptr, nullptr, nullptr, nullptr, nullptr);
        respond(S, a, out, rc);
        if (rc == E3X_OK) retire_handle(S, a.hd);
        else drop_handle(S, a.hd);
        S.active.erase(S.active.begin() + (std::ptrdiff_t)i);
        *changed = true;
    }
    // admit queued requests: first fit in arrival order; a request of a small grid takes
    // the queued requests it can share a batch with (in arrival order) along.  With
    // small_slots > 0 the solves of small grids have that many slots of their own and the
    // others max_active; else all share max_active.
    auto small_grid = [&](const e3x_request &q) { return (int64_t)q.nx * q.ny * q.nz <= S.batch_max_nodes; };
    auto slot_free = [&](const e3x_request &q) {
        if (S.small_slots <= 0) return (int)S.active.size() < S.max_active;
        int nsmall = 0;
        for (const Active &a : S.active) nsmall += a.small ? 1 : 0;
        return small_grid(q) ? nsmall < S.small_slots : (int)S.active.size() - nsmall < S.max_active;
    };
    for (auto it = S.queue.begin(); it != S.queue.end();) {
        if (S.small_slots <= 0 && (int)S.active.size() >= S.max_active) break;
        if (!slot_free(it->rq)) {
            ++it;
            continue;
        }
        const int ps = prep_state(*it);
        if (ps == 0) {  // not prepared yet
            ++it;
            continue;
        }
        if (ps < 0) {
            e3x_response o = make_response(*it, report_error(E3X_EPHASE, "phase label in pix out of bounds"));
            o.t_admit = o.t_done = S.now();
            S.done.push_back(o);
            it = S.queue.erase(it);
            *changed = true;
            continue;
        }
        const e3x_request &q = it->rq;
        std::vector<std::list<Request>::iterator> group{it};
        if (S.batch_members > 1 && (int64_t)q.nx * q.ny * q.nz <= S.batch_max_nodes)
            for (auto jt = std::next(it); jt != S.queue.end() && (int)group.size() < S.batch_members; ++jt)
                if (same_batch(jt->rq, q) && prep_state(*jt) == 1) group.push_back(jt);
        const int members = group.size() > 1 ? (int)group.size() : 0;
        int np_each = 0;
        for (auto &g : group) np_each = std::max(np_each, g->rq.nphase);
        bool fields = false;  // the full-field staging a handle allocates on first use
        for (auto &g : group) fields |= g->rq.stress_field != nullptr || g->rq.vm != nullptr;
        const int64_t need = footprint(q.nx, q.ny, q.nz, std::max(kPoolPhases, np_each),
                                       q.precision == E3X_PREC_MIXED, q.ldemb) *
                                 std::max(members, 1) +
                             (fields ? 7 * 8 * (int64_t)q.nx * q.ny * q.nz : 0);
        Handle hd;
        bool fits = true;
        if (int rc = acquire_handle(S, q, members, np_each, fields, need, &hd, &fits)) {
            e3x_response o = make_response(*it, rc);
            o.t_admit = o.t_done = S.now();
            S.done.push_back(o);
            it = S.queue.erase(it);
            *changed = true;
            continue;
        }
        if (!fits) {
            ++it;
            continue;
        }
        Active a;
        a.hd = hd;
        a.small = small_grid(q);
        a.t_admit = S.now();
        for (auto &g : group) a.reqs.push_back(std::move(*g));
        for (size_t k = 1; k < group.size(); k++) S.queue.erase(group[k]);  // never the current one
        it = S.queue.erase(it);
        *changed = true;
        if (int rc = start_solve(a)) {
            std::vector<e3x_response> out;
            for (const Request &r : a.reqs) out.push_back(make_response(r, rc));
            respond(S, a, out, rc);
            if (reinterpret_cast<Solver *>(a.hd.h)->run.on)
                (void)e3x_dembx_end(a.hd.h, nullptr, nullptr, nullptr, nullptr, nullptr);
            // an input error leaves the handle sound; a device error may not
            if (rc == E3X_EPHASE || rc == E3X_EINVAL) retire_handle(S, a.hd);
            else drop_handle(S, a.hd);
            continue;
        }
        S.active.push_back(std::move(a));
    }
    // keep every active PCG supplied with iteration chunks
    for (Active &a : S.active) {
        int fin = 0;
        (void)e3x_dembx_advance(a.hd.h, 0, &fin);
    }
    return E3X_OK;
}

}  // namespace

extern "C" {

int e3x_server_create(int device, int64_t hbm_budget_bytes, int max_active, e3x_server **out)
{
    if (!out) return report_error(E3X_EINVAL, "out is NULL");
    *out = nullptr;
    if (max_active < 1 || max_active > 64) return report_error(E3X_EINVAL, "max_active must be in [1, 64]");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n)
        return report_error(E3X_EINVAL, "no such HIP device");
    if (hipSetDevice(device) != hipSuccess) return report_error(E3X_EDEVICE, "hipSetDevice failed");
    int64_t budget = hbm_budget_bytes;
    if (budget <= 0) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) return report_error(E3X_EDEVICE, "hipMemGetInfo failed");
        budget = (int64_t)(0.8 * (double)fr);
    }
    Server *s = new (std::nothrow) Server;
    if (!s) return report_error(E3X_ENOMEM, "host allocation failed");
    s->device = device;
    s->max_active = max_active;
    s->budget = budget;
    s->t0 = std::chrono::steady_clock::now();
    *out = reinterpret_cast<e3x_server *>(s);
    return E3X_OK;
}

int e3x_server_destroy(e3x_server *sv)
{
    if (!sv) return E3X_OK;
    Server &S = SV(sv);
    for (Active &a : S.active) {
        (void)e3x_dembx_end(a.hd.h, nullptr, nullptr, nullptr, nullptr, nullptr);
        (void)e3x_destroy(a.hd.h);
    }
    for (Handle &hd : S.pool) (void)e3x_destroy(hd.h);
    delete &S;
    return E3X_OK;
}

int e3x_server_submit(e3x_server *sv, const e3x_request *rq, int64_t *id)
{
    if (!sv || !rq) return report_error(E3X_EINVAL, "NULL server or request");
    Server &S = SV(sv);
    if (rq->nx < 2 || rq->ny < 2 || rq->nz < 2) return report_error(E3X_EINVAL, "grid extents must be >= 2");
    if ((int64_t)rq->nx * rq->ny * rq->nz > (int64_t)INT32_MAX) return report_error(E3X_EINVAL, "grid too large");
    if (rq->nphase < 1 || rq->nphase > kMaxPhases) return report_error(E3X_EINVAL, "nphase must be in [1, 2^31 - 2]");
    if (!rq->pix) return report_error(E3X_EINVAL, "pix is NULL");
    if (rq->flag_m != 0 && rq->nphase > 1 && !rq->orientation)
        return report_error(E3X_EINVAL, "orientation is NULL");
    const int64_t need = footprint(rq->nx, rq->ny, rq->nz, std::max(kPoolPhases, rq->nphase),
                                   rq->precision == E3X_PREC_MIXED, rq->ldemb);
    if (need > S.budget) return report_error(E3X_ENOMEM, "request exceeds the server's memory budget");
    Request r;
    r.id = S.next_id++;
    r.rq = *rq;
    const int64_t ns = (int64_t)rq->nx * rq->ny * rq->nz;
    r.prep = std::make_shared<Prep>();
    r.prep->pix.assign(rq->pix, rq->pix + ns);
    r.prep->nphase = rq->nphase;
    if (rq->orientation && rq->nphase > 1) r.ori.assign(rq->orientation, rq->orientation + 3 * (rq->nphase - 1));
    r.rq.pix = nullptr;
    r.rq.orientation = nullptr;
    r.t_submit = S.now();
    if (ns <= kPrepInline) prepare(*r.prep);  // ready at once (under ~0.3 ms of host work)
    else S.prep.push(r.prep);
    S.queue.push_back(std::move(r));
    if (id) *id = S.next_id - 1;
    return E3X_OK;
}

int e3x_server_progress(e3x_server *sv, int *n_active, int *n_queued, int *n_done)
{
    if (!sv) return report_error(E3X_EINVAL, "NULL server");
    Server &S = SV(sv);
    if (hipSetDevice(S.device) != hipSuccess) return report_error(E3X_EDEVICE, "hipSetDevice failed");
    bool changed = false;
    if (int rc = progress(S, &changed)) return rc;
    if (n_active) *n_active = (int)S.active.size();
    if (n_queued) *n_queued = (int)S.queue.size();
    if (n_done) *n_done = (int)S.done.size();
    return E3X_OK;
}

int e3x_server_take(e3x_server *sv, e3x_response *out, int *got)
{
    if (!sv || !out || !got) return report_error(E3X_EINVAL, "NULL argument");
    Server &S = SV(sv);
    *got = 0;
    if (S.done.empty()) return E3X_OK;
    *out = S.done.front();
    S.done.pop_front();
    *got = 1;
    return E3X_OK;
}

int e3x_server_drain(e3x_server *sv)
{
    if (!sv) return report_error(E3X_EINVAL, "NULL server");
    Server &S = SV(sv);
    if (hipSetDevice(S.device) != hipSuccess) return report_error(E3X_EDEVICE, "hipSetDevice failed");
    while (!S.queue.empty() || !S.active.empty()) {
        bool changed = false;
        if (int rc = progress(S, &changed)) return rc;
        if (!changed) std::this_thread::sleep_for(std::chrono::microseconds(100));
        if (!changed && S.active.empty() && !S.queue.empty())
            return report_error(E3X_ENOMEM, "queued requests cannot be admitted under the memory budget");
    }
    return E3X_OK;
}

int e3x_server_set_batching(e3x_server *sv, int max_members, int64_t max_nodes)
{
    if (!sv) return report_error(E3X_EINVAL, "NULL server");
    if (max_members < 1 || max_members > kMaxRhs) return report_error(E3X_EINVAL, "max_members must be in [1, 6]");
    if (max_nodes < 0) return report_error(E3X_EINVAL, "max_nodes must be >= 0");
    Server &S = SV(sv);
    S.batch_members = max_members;
    S.batch_max_nodes = max_nodes;
    return E3X_OK;
}

int e3x_server_set_small_slots(e3x_server *sv, int slots)
{
    if (!sv) return report_error(E3X_EINVAL, "NULL server");
    if (slots < 0 || slots > 64) return report_error(E3X_EINVAL, "slots must be in [0, 64]");
    SV(sv).small_slots = slots;
    return E3X_OK;
}

int e3x_server_clock(const e3x_server *sv, double *seconds)
{
    if (!sv || !seconds) return report_error(E3X_EINVAL, "NULL argument");
    *seconds = SV(sv).now();
    return E3X_OK;
}

int e3x_server_memory(const e3x_server *sv, int64_t *bytes, int *handles)
{
    if (!sv) return report_error(E3X_EINVAL, "NULL server");
    const Server &S = SV(sv);
    if (bytes) *bytes = S.used;
    if (handles) *handles = (int)(S.active.size() + S.pool.size());
    return E3X_OK;
}

}  // extern "C"
