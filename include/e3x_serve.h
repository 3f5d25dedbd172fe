/*
 * e3x_serve.h -- request loop over the gfx950 solver (no counterpart in the reference,
 * which runs one solve per process).
 *
 * A server owns one GPU.  Requests (a microstructure, constants, one macro strain and
 * the PCG controls) are queued on submit; every call of e3x_server_progress retires
 * the solves that have finished, admits queued requests while the active count stays
 * under `max_active` and the device memory of the server's solver handles under
 * `hbm_budget_bytes` (first fit in arrival order: a request that does not fit yet lets
 * later, smaller ones pass; requests on one small grid are admitted together as a batch,
 * e3x_server_set_batching), and keeps the PCG of every active solve running on its own
 * stream.  Solves of different grid sizes therefore run side by side on the GPU,
 * and a request that arrives while others run starts as soon as a slot and memory are
 * free.  Each admitted request runs the main program's sequence (f90:380-519: ppixel,
 * femat, affine guess, energy, dembx, energy, stress) with its state resident on the
 * device; finished handles are kept for later requests of the same grid size and freed
 * only when memory is needed.  Results do not depend on what else runs: every kernel
 * sums in a fixed order.
 *
 * The host work on a request's labels (the check against nphase, the phase histogram,
 * the conversion to the device's phase ids) is done at submit for grids of at most 64^3
 * nodes, and otherwise by a preparation thread of the server; progress admits a request
 * once it is prepared.  So the thread that drives the GPU does not stop for it (a 200^3
 * request is ~10 ms of it) while other solves wait for their next iterations.
 *
 * One host thread drives a server (submit, progress, take, drain are not thread-safe
 * among themselves); different servers are independent.  Every call returns E3X_OK or
 * an e3x_status with e3x_last_error() set.
 */
#ifndef E3X_SERVE_H
#define E3X_SERVE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct e3x_server e3x_server;

typedef struct e3x_request {
    int nx, ny, nz, nphase;
    const int32_t *pix;        /* [nx*ny*nz], labels 1..nphase (copied on submit)              */
    const double *orientation; /* [3*(nphase-1)] Rodrigues vectors (copied); NULL for flag_m 0 */
    double c11, c12, c44;
    double phasemod[4];        /* as e3x_femat                                                  */
    int flag_m;
    double strain6[6];         /* {exx, eyy, ezz, exz, eyz, exy} tensor components             */
    int precision;             /* E3X_PREC_FP64 or E3X_PREC_MIXED                               */
    int replace_every;         /* mixed form: true-residual replacement period (0: never)       */
    float tol;                 /* stop tolerance, float32(1e-8) in the reference                */
    int ldemb;                 /* iteration limit                                               */
    double *stress_field;      /* optional full-field response (may be NULL): host buffers the */
    double *vm;                /* server fills before the response is ready, stress (6,ns) as  */
                               /* Fortran stress_field(ns,6) and von Mises (ns), as            */
                               /* stress_fullfield_OpenMP computes them (f90:2639-2833); owned */
                               /* by the caller until the response is taken                    */
} e3x_request;

typedef struct e3x_response {
    int64_t id;
    int status;                /* E3X_OK, or the status of the step that failed                */
    int lstep, exit_flag, reset_count;
    int nx, ny, nz;
    double utot0, utot, gg;    /* energies before / after the PCG, final z.g                    */
    double stress6[6], strain6[6];  /* averages (xx yy zz xz yz xy; engineering strains)       */
    double t_submit, t_admit, t_done;  /* seconds on the server's clock                         */
} e3x_response;

/*
 * hbm_budget_bytes <= 0: 80 % of the device memory free at creation.  max_active: solves
 * running at once (1..64).
 */
int e3x_server_create(int device, int64_t hbm_budget_bytes, int max_active, e3x_server **out);

/* Wait for the active solves, free every handle and the server. */
int e3x_server_destroy(e3x_server *sv);

/* Queue a request (inputs are copied); *id numbers requests from 0 in submit order. */
int e3x_server_submit(e3x_server *sv, const e3x_request *rq, int64_t *id);

/* One scheduling pass (retire, admit, advance); does not wait for the GPU.  Any of the
 * counters may be NULL. */
int e3x_server_progress(e3x_server *sv, int *n_active, int *n_queued, int *n_done);

/* Pop one finished response (oldest first): *got = 1 and *out filled, or *got = 0. */
int e3x_server_take(e3x_server *sv, e3x_response *out, int *got);

/* Run progress until nothing is queued or active (waits on the GPU between passes). */
int e3x_server_drain(e3x_server *sv);

/*
 * Co-batching: queued requests on one grid of at most max_nodes nodes, with the same
 * precision and PCG controls (replace_every, tol, ldemb, flag_m), are admitted together,
 * up to max_members of them (1..6; 1 turns co-batching off), as the members of one batch
 * handle (e3x_create_batch): their problems share every launch.  A batch takes one slot of
 * max_active and its requests finish together; each result is bit for bit the one its
 * request gets alone.  Defaults: 6 members, 128^3 nodes.
 */
int e3x_server_set_batching(e3x_server *sv, int max_members, int64_t max_nodes);

/*
 * Slots of their own for the solves of small grids (at most the co-batching max_nodes):
 * with slots > 0 they take up to `slots` slots beside the max_active ones, which the
 * other solves keep for themselves, so a small request does not queue behind large
 * solves.  0 (the default): every solve shares max_active.
 */
int e3x_server_set_small_slots(e3x_server *sv, int slots);

/* Seconds since the server was created (the clock of t_submit, t_admit, t_done). */
int e3x_server_clock(const e3x_server *sv, double *seconds);

/* Device bytes held by the server's handles (active and pooled), and the handle count. */
int e3x_server_memory(const e3x_server *sv, int64_t *bytes, int *handles);

#ifdef __cplusplus
}
#endif
#endif /* E3X_SERVE_H */
