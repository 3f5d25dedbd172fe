/*
 * e3x.h -- C ABI of the MI355X (gfx950) elas3dxtal_pcg solver.
 *
 * Drop-in boundary for the reference's solver routines in
 * src/elas3dxtal_pcg.f90.  Each entry point below replaces one subroutine (cited)
 * and is callable from Fortran through iso_c_binding (plain pointers, sizes and
 * scalars; arrays in Fortran column-major layout, unchanged).  All device state
 * lives in an opaque handle; NULL host pointers mean "keep it on the device".
 *
 * Conventions (all 0-based in C, identical memory layout to the Fortran arrays):
 *   grid      nx*ny*nz periodic nodes, node m = i + nx*(j + ny*k)   (f90:350)
 *   pix       int32[ns], phase id 1..nphase of the voxel at node m (f90:2945)
 *   vectors   double[nrhs][3][ns]  == Fortran u(ns,3) per right-hand side
 *   strain6   {exx, eyy, ezz, exz, eyz, exy} tensor components       (f90:394-399)
 *   stress6   averages in order xx, yy, zz, xz, yz, xy; strains engineering
 *             (f90:2379-2391)
 *
 * Every call returns E3X_OK (0) or a nonzero e3x_status; e3x_last_error() then
 * holds a message (per thread).  Calls on one handle are synchronous and must
 * come from one host thread at a time; different handles are independent.
 */
#ifndef E3X_H
#define E3X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct e3x_solver e3x_solver;

enum e3x_status {
    E3X_OK = 0,
    E3X_EINVAL = 1,   /* bad argument (sizes, NULL where required, ranges)          */
    E3X_EDEVICE = 2,  /* HIP runtime error                                          */
    E3X_ENOMEM = 3,   /* device allocation failed                                   */
    E3X_EPHASE = 4,   /* pix label outside [1, nphase]   (replaces f90:2986-2989 stop) */
    E3X_ESTATE = 5,   /* call order violated (e.g. dembx before femat)              */
    E3X_EIO = 6       /* file I/O (HDF5 helpers)                                    */
};

/* How a PCG solve ended (dembx_OpenMP, f90:1899-1902, 2007, loop end). */
enum e3x_pcg_exit {
    E3X_PCG_RUNNING = 0,
    E3X_PCG_CONVERGED = 1, /* rel_res < tol (f90:2007)                       */
    E3X_PCG_HAH_BREAK = 2, /* |h.Ah| < 1e-14: WARNING and break (f90:1899)   */
    E3X_PCG_MAXITER = 3    /* ldemb iterations without convergence          */
};

/* PCG stop rule (e3x_set_stop_rule). */
enum e3x_stop_rule {
    E3X_STOP_REFERENCE = 0, /* sqrt(z.g / gg0) < tol, z = M^-1 g (f90:1949, 2007): default */
    E3X_STOP_NIST_ABS = 1   /* sum of g^2 < 1e-8 * ns on the unpreconditioned gradient:
                               NIST ELAS3D's absolute rule, the wording of the reference's
                               README; meaningful for moduli of order one (GPa units),
                               not for moduli in Pa */
};

/* PCG loop variant (e3x_set_pcg_variant): the loop of one of the reference's programs,
 * all on the same stencil, reductions and stop rule. */
enum e3x_pcg_variant {
    E3X_PCG_REFERENCE = 0, /* src/elas3dxtal_pcg.f90 dembx_OpenMP (f90:1356-2043): block-
                              Jacobi preconditioner, 10x divergence reset (f90:1961-1981),
                              direction update before the stop test (f90:1996, 2007).
                              Default. */
    E3X_PCG_SWEEP = 1,     /* applications/sec3_2/A_x_y/elas3dxtal_pcg.f90 dembx_OpenMP
                              (A_2_0:1045-1558): block Jacobi, no reset, stop test before
                              the direction update (A_2_0:1518-1526) */
    E3X_PCG_PLAIN_CG = 2   /* applications/sec3_1/cg/n_N/elas3dxtal_cg.f90 dembx_OpenMP
                              (cg:915-1074; the serial variant's loop is the same): no
                              preconditioner (z = g, h = gb at entry, cg:949), no reset,
                              stop test before the direction update (cg:1050-1056) */
};

/* Fields reachable with e3x_get_field / e3x_set_field. */
enum e3x_field {
    E3X_FIELD_U = 0,  /* displacement u(ns,3)          (e3d%u)  */
    E3X_FIELD_GB = 1, /* gradient gb = K u + b (ns,3)  (e3d%gb) */
    E3X_FIELD_B = 2,  /* right-hand side b (ns,3)      (e3d%b)  */
    E3X_FIELD_H = 3   /* search direction h (ns,3)     (e3d%h)  */
};

/* Arithmetic of e3x_dembx_ex. */
enum e3x_precision {
    E3X_PREC_FP64 = 0,  /* the reference's arithmetic                                   */
    E3X_PREC_MIXED = 1  /* FP32 stencil, vectors and block-Jacobi, the u steps in an FP32
                           increment folded into the FP64 u at each true-residual
                           replacement (every `replace_every` steps) and at the end; FP64
                           dots, norms and replacement                                  */
};

/* Message of the last failing call on this thread ("" if none). */
const char *e3x_last_error(void);

/* Library ABI version (major*10000 + minor*100 + patch). */
int e3x_version(void);

/* Number of visible HIP devices. */
int e3x_device_count(int *count);

/*
 * Create a solver for an nx*ny*nz periodic grid with nphase phases (grains plus
 * the pore phase) and nrhs simultaneous right-hand sides (1..6; 6 = the unit
 * load cases of the effective stiffness).  Replaces the allocations of
 * program elas3d (f90:242-256).  nphase <= 2^31 - 2 (phase ids are held in 16 bits up to
 * 65,535 phases and in 32 bits above, as the reference's anisotropy-sweep inputs with
 * 81,488 grains need); every extent >= 2 (E3X_EINVAL otherwise: a one-node-thick
 * periodic direction is not supported).
 */
int e3x_create(int nx, int ny, int nz, int nphase, int nrhs, int device, e3x_solver **out);

/*
 * A batch of `members` (2..6) separate problems on one nx*ny*nz grid, solved together.  No
 * counterpart in the reference, which solves one problem per run: it is how the serving
 * loop co-batches small grids.  Right-hand side r of the handle is member r, with its own
 * microstructure (nphase[r] phases), materials and load, set by e3x_set_member; the
 * members share every launch (the stencil grid has a row per member, the streaming
 * kernels take them in turn), so a few small grids fill the GPU.  Once every member is
 * set, e3x_affine_guess, e3x_energy, e3x_dembx(_ex) / _begin / _advance / _end,
 * e3x_stress, e3x_stress_fullfield, e3x_hotspots, e3x_get_field and e3x_set_field take
 * the members as they take right-hand sides, and each member's results are bit for bit
 * those of its problem solved alone (the same kernels, grids and order of every sum).
 * e3x_set_microstructure, e3x_assig, e3x_femat and e3x_effective_stiffness are not for
 * batch handles (E3X_ESTATE).
 */
int e3x_create_batch(int nx, int ny, int nz, int members, const int *nphase, int device, e3x_solver **out);

/*
 * Member r of a batch: its microstructure (pix labels in [1, nphase[r]] and Rodrigues
 * orientation[3*(nphase[r]-1)], as e3x_set_microstructure), materials and load (as
 * e3x_femat with one strain6).  Builds the member's tables and lane order; the call that
 * sets the last member also builds every member's jump right-hand side and constant
 * (f90:815-1100).  Setting a member again replaces it (the others are kept).
 */
int e3x_set_member(e3x_solver *s, int member, const int32_t *pix, const double *orientation, double c11,
                   double c12, double c44, const double *phasemod, int flag_m, const double *strain6);

/* Free all device memory of the handle (f90:543-549). */
int e3x_destroy(e3x_solver *s);

/* Bytes of device memory the handle holds now. */
int e3x_memory_bytes(const e3x_solver *s, int64_t *bytes);

/*
 * Load the microstructure: pix[ns] with labels in [1, nphase] (checked,
 * E3X_EPHASE otherwise) and orientation[3*(nphase-1)] Rodrigues vectors
 * tan(theta/2)*axis, grain-major.  Data part of ppixel_hdf5 (f90:2879-2991);
 * the HDF5 file reading itself is in e3x_io.h.  orientation may be NULL for
 * flag_m = 0.
 */
int e3x_set_microstructure(e3x_solver *s, const int32_t *pix, const double *orientation);

/* Phase volume fractions prob[nphase] (assig, f90:2410-2468). */
int e3x_assig(e3x_solver *s, double *prob);

/*
 * femat_OpenMP (f90:555-1104): per-phase stiffness cmod (rotated cubic c11,c12,c44
 * per grain for flag_m = 1, rotatestiffnessby f90:2470-2637; isotropic K,G for
 * flag_m = 0), element blocks dk, the per-phase 27-point stencil blocks, and for
 * each right-hand side r the periodic-jump force b_r and constant C_r from
 * strain6[6*r .. 6*r+5].
 *   phasemod: flag_m = 1 -> {K_pore, G_pore} (2 values, f90:301-317);
 *             flag_m = 0 -> {K1, G1, K2, G2}  (f90:295-312, nphase must be 2).
 *   C_out[nrhs] (may be NULL).
 */
int e3x_femat(e3x_solver *s, double c11, double c12, double c44, const double *phasemod,
              int flag_m, const double *strain6, double *C_out);

/* Affine initial guess u = E x for every right-hand side (f90:423-438). */
int e3x_affine_guess(e3x_solver *s);

/*
 * energy_OpenMP (f90:1107-1353): gb = K u + b and utot = C + sum(u.Ku/2 + b.u)
 * for every right-hand side; also gg[r] = sum(gb^2) (f90:452).  utot[nrhs],
 * gg[nrhs] may be NULL.
 */
int e3x_energy(e3x_solver *s, double *utot, double *gg);

/*
 * dembx_OpenMP (f90:1356-2043): block-Jacobi PCG from the current u, gb.
 *   gg0[nrhs]    initial sum(gb^2) per right-hand side (f90:452-453)
 *   tol          single-precision tolerance on sqrt(z.g/gg0) (f90:148, 1949, 2007)
 *   ldemb        iteration limit (f90:141)
 *   kkk          outer step; kkk == 1 builds the preconditioner and sets h = z
 *                (f90:1498, 1554); kkk > 1 keeps both (the reference reuses h and
 *                leaves M_block_inv unset, f90:1469)
 *   lstep[nrhs], gg[nrhs]           iterations taken, final z.g
 *   history[nrhs*ldemb]             rel_res per iteration (f90:1952-1953, 2020); may be NULL
 *   exit_flag[nrhs], reset_count[nrhs]  e3x_pcg_exit and divergence resets
 *                                   (f90:1961-1981); may be NULL
 */
int e3x_dembx(e3x_solver *s, const double *gg0, float tol, int ldemb, int kkk, int *lstep,
              double *gg, double *history, int *exit_flag, int *reset_count);

/*
 * e3x_dembx with a precision mode.  E3X_PREC_FP64 is e3x_dembx.  E3X_PREC_MIXED
 * runs the stencil, vectors and preconditioner in FP32, keeps u, the dot products
 * and norms in FP64, and every `replace_every` iterations replaces the recursive
 * gradient by the true one, g = K u + b, computed in FP64 (reliable update).
 * No counterpart in the reference (its solver is FP64 only).
 */
int e3x_dembx_ex(e3x_solver *s, int precision, int replace_every, const double *gg0, float tol,
                 int ldemb, int kkk, int *lstep, double *gg, double *history, int *exit_flag,
                 int *reset_count);

/*
 * Stop rule of the handle's following e3x_dembx / e3x_dembx_ex / e3x_dembx_begin calls
 * (e3x_stop_rule; E3X_STOP_REFERENCE until set).  With E3X_STOP_NIST_ABS the relative
 * residual is still recorded in the history and `tol` is not used.  No counterpart in
 * the reference, whose code has only the relative rule (f90:2007).
 */
int e3x_set_stop_rule(e3x_solver *s, int rule);

/*
 * PCG loop of the handle's following e3x_dembx / e3x_dembx_ex / e3x_dembx_begin calls
 * (e3x_pcg_variant; E3X_PCG_REFERENCE until set): dembx_OpenMP of src/elas3dxtal_pcg.f90
 * (f90:1356), of the anisotropy-sweep programs (A_2_0:1045) or of the plain-CG programs
 * (cg:915).  With E3X_PCG_PLAIN_CG the preconditioner is the identity.
 */
int e3x_set_pcg_variant(e3x_solver *s, int variant);

/*
 * Hotspots of a von Mises field as the reference's MATLAB post-processing finds them for
 * top_peaks.txt (applications/sec3_2/A_2_0/elas3dxtal_postprocessing.m:86-243; the sec3_3
 * scripts take the same steps):
 *   1. solid voxels: phase label (1-based) <= nsolid.  The scripts set nsolid = max(pix)
 *      (.m:67), which counts pores as solid;
 *   2. masked normalised Gaussian smoothing with zero padding (the grid is not wrapped):
 *      s = G*(vm on the solid) / G*(solid), G = exp(-d^2 / (2 sigma^2)) for |d| <=
 *      ceil(2 sigma) per axis, normalised (imgaussfilt3).  The numerator is double; the
 *      denominator and the quotient are single precision, as the script's double ./
 *      single gives.  A denominator below 1e-6 counts as 1.  s = 0 off the solid
 *      (.m:86-97);
 *   3. regional maxima of s with 26-connectivity (imregionalmax, .m:206): connected sets
 *      of equal value whose outside neighbours are all lower (every member counts);
 *   4. those with s / 1e6 >= threshold_mpa, in single precision (.m:207-208), sorted by
 *      value, descending, ties by linear index (.m:215).
 * vm: ns values in node order (the /vm dataset of fullfield_poly.h5), or NULL for the
 * handle's own full field of right-hand side rhs (computed on the device as
 * e3x_stress_fullfield does; needs e3x_femat).  Writes the first npeaks (<= 64) peaks:
 * index[k] = i + nx (j + ny k), 0-based, and value[k] = s there in Pa (single precision
 * widened), -1 and 0 past the last one.  *found (may be NULL) is the number of regional
 * maxima at or above the threshold.  sigma in (0, 1.5].  No counterpart in the
 * reference's Fortran: it is its MATLAB post-processing, on the device.
 */
int e3x_hotspots(e3x_solver *s, int rhs, const double *vm, int nsolid, double sigma, double threshold_mpa,
                 int npeaks, int64_t *index, double *value, int64_t *found);

/*
 * e3x_dembx_ex in three steps, so that one host thread can keep several handles'
 * solves in flight on the GPU at once (each handle has its own stream); no counterpart
 * in the reference.  e3x_dembx_begin validates, builds the preconditioner and enqueues
 * the PCG entry; e3x_dembx_advance enqueues further iteration chunks and sets *done = 1
 * once the loop and its final updates have completed on the device (block = 0: returns
 * at once if the GPU is still working; block = 1: waits); e3x_dembx_end returns what
 * e3x_dembx_ex returns (it waits for completion if needed).  No other call on the handle
 * between begin and end.
 */
int e3x_dembx_begin(e3x_solver *s, int precision, int replace_every, const double *gg0, float tol, int ldemb,
                    int kkk);
int e3x_dembx_advance(e3x_solver *s, int block, int *done);
int e3x_dembx_end(e3x_solver *s, int *lstep, double *gg, double *history, int *exit_flag, int *reset_count);

/* stress_OpenMP (f90:2149-2408): stress6[6*nrhs], strain6[6*nrhs] volume averages. */
int e3x_stress(e3x_solver *s, double *stress6, double *strain6);

/*
 * e3x_stress on caller-owned device memory, allocating nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops that forbid hipMalloc).  Enqueues on `stream` (a hipStream_t;
 * NULL: the null stream) of the calling thread's current device the averages of
 * stress_OpenMP (f90:2149-2408) and returns without waiting:
 *   pix       device int32[ns], labels 1..nphase (not checked on the host; a voxel with a
 *             label outside [1, nphase] makes the averages NaN)
 *   cmod      device double[nphase*36] (cmod(p,i,j) row-major per phase, e3x_get_tables)
 *   strain6   host double[6*nrhs], the applied strains as e3x_femat takes them
 *   u         device double[nrhs*3*ns], u(ns,3) per right-hand side (component planes)
 *   averages  device double[nrhs*12]: per right-hand side stress6 then strain6, as
 *             e3x_stress returns them (the same kernel and summation order)
 * The block partials go to a scratch array of 12*6*2048 doubles (1.2 MB) that the library
 * holds on each GPU it runs on until the process ends: nrhs is at most 6 (E3X_EINVAL
 * otherwise).  Every call on one GPU shares that GPU's scratch: calls that overlap on one
 * GPU (other streams, other host threads) race, so serialise them.
 */
int e3x_stress_dev(int nx, int ny, int nz, int nphase, int nrhs, const int32_t *pix, const double *cmod,
                   const double *strain6, const double *u, double *averages, void *stream);

/*
 * stress_fullfield_OpenMP computation part (f90:2639-2833) for right-hand side
 * `rhs`: stress[6][ns] (Fortran stress_field(ns,6)) and von Mises vm[ns]; either
 * may be NULL.  The HDF5 write (f90:2840-2874) is e3x_h5_write_fullfield.
 */
int e3x_stress_fullfield(e3x_solver *s, int rhs, double *stress, double *vm);

/*
 * Effective stiffness of the loaded microstructure (no counterpart in the reference,
 * which solves one load case per run): on a handle created with nrhs = 6, runs the main
 * program's sequence (f90:417-514: femat, affine guess, energy, dembx, stress) for the six
 * unit macro strains of size `magnitude` (3 normal; 3 engineering shears xz, yz, xy,
 * tensor component magnitude/2) as six right-hand sides of one batched PCG, and returns
 * C_eff[36] (row-major, C_eff[6*i + j] = d<sigma_i>/d<eps_j>, Voigt order xx yy zz xz
 * yz xy, engineering shear strains) from the averaged stresses and strains:
 * C_eff = S E^-1.  precision / replace_every as in e3x_dembx_ex; lstep[6], exit_flag[6]
 * may be NULL.
 */
int e3x_effective_stiffness(e3x_solver *s, double c11, double c12, double c44, const double *phasemod,
                            int flag_m, double magnitude, int precision, int replace_every, float tol,
                            int ldemb, double *C_eff, int *lstep, int *exit_flag);

/* Copy a field (enum e3x_field) of right-hand side `rhs` to / from host double[3*ns]. */
int e3x_get_field(e3x_solver *s, int which, int rhs, double *host);
int e3x_set_field(e3x_solver *s, int which, int rhs, const double *host);

/* Copy the phase tables to the host: cmod[nphase*36] (cmod(p,i,j) row-major per
 * phase) and dk[nphase*576] (dk(p,a,i,b,j) with j fastest); either may be NULL. */
int e3x_get_tables(e3x_solver *s, double *cmod, double *dk);

/* GPU time of the last e3x_dembx(_ex) call: preconditioner build and the PCG loop (ms). */
int e3x_last_timing(e3x_solver *s, double *precond_ms, double *pcg_ms);

/*
 * Measurement hook (no counterpart in the reference): with on != 0, the handle's
 * following e3x_dembx / e3x_dembx_ex / e3x_dembx_begin runs record HIP events on the
 * solver's stream around the stencil-apply launch of every iteration from the second on
 * (the fused form; at most the first 1024 of a run), and e3x_last_apply_timing returns
 * how many of those launches did work (iterations up to the last stop) and their mean
 * duration in ms: the apply as the PCG loop runs it, between its neighbours in the loop.
 */
int e3x_set_apply_timing(e3x_solver *s, int on);
int e3x_last_apply_timing(e3x_solver *s, int *launches, double *avg_ms);

/*
 * Launch tuning of the PCG's stencil apply (no counterpart in the reference): the waves per
 * SIMD its register budget is compiled for, FP64 {5, 6, 8} and FP32 {6, 8}.  With waves = 0
 * (the default) the first e3x_dembx* (or e3x_time_apply) of a precision times every
 * variant on the handle's own data (four launches each) and keeps the fastest; the
 * choice is then shared by the process's handles of the same grid, right-hand sides and
 * precision.  waves != 0 forces that variant on this handle.  Every variant computes the
 * same bits.  e3x_get_tuning returns the variant in use (0 before the first PCG) and its
 * measured mean apply time in ms (0 when forced or not measured).
 */
int e3x_set_tuning(e3x_solver *s, int precision, int waves);
int e3x_get_tuning(e3x_solver *s, int precision, int *waves, double *apply_ms);

/*
 * Measurement hook (no counterpart in the reference): replay `reps` launches of the
 * PCG's stencil-apply kernel in the form the loop runs from its second iteration on
 * (direction h = z + gamma h_prev, u -= lambda h_prev, Ah = K h with the h.Ah
 * partials; FP64 or the mixed form's FP32) on the solver's stream, bracketed by HIP
 * events on that stream, and return the average launch time in ms.  It runs with
 * lambda = gamma = 0, so u keeps its values; the PCG state is saved and restored; the
 * second direction buffer, Ah and the partials are overwritten.  Needs e3x_femat first.
 */
int e3x_time_apply(e3x_solver *s, int precision, int reps, double *avg_ms);

#ifdef __cplusplus
}
#endif
#endif /* E3X_H */
