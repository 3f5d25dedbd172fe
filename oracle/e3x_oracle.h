/*
 * e3x_oracle.h -- CPU restatement of the elas3dxtal_pcg solve (TEST INFRASTRUCTURE).
 *
 * This is the checker, not the product: only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it.  The product path is the HIP library
 * declared in include/e3x.h and never calls into this code.
 *
 * Every function restates one routine of src/elas3dxtal_pcg.f90 (cited per
 * function below).  Layouts follow the reference with 0-based C indices:
 *   node m = i + nx*(j + ny*k)                     (f90:350, i fastest)
 *   vectors  v[c*ns + m]  (c = 0,1,2)              (Fortran v(ns,3), column-major)
 *   pix[m]   phase id 1..nphase of the voxel whose lower corner is node m (f90:2945)
 *   cmod[(p*6 + r)*6 + c]                          (cmod(p,r,c))
 *   dk[(((p*8 + a)*3 + i)*8 + b)*3 + j]            (dk(p,a,i,b,j))
 *   strain6 = {exx, eyy, ezz, exz, eyz, exy}       tensor components (f90:167, 394-399)
 */
#ifndef E3X_ORACLE_H
#define E3X_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* Result flags of oc_dembx (f90:1899-1902, 1961-1981, 2007). */
enum { OC_RUNNING = 0, OC_CONVERGED = 1, OC_HAH_BREAK = 2, OC_MAXITER = 3 };

/* f90:326-369: periodic 27-neighbour table, ib[n*ns + m] (0-based node ids). */
void oc_build_ib(int nx, int ny, int nz, int32_t *ib);

/* f90:2410-2468: phase volume fractions prob[nphase]. */
void oc_assig(int64_t ns, int nphase, const int32_t *pix, double *prob);

/* f90:2470-2637: Rodrigues -> Q -> Bond matrix -> rot = newr^T C newr, rot36[g*36 + r*6 + c]. */
void oc_rotate_stiffness(int ngrain, const double *rod, double c11, double c12, double c44,
                         double *rot36);

/* f90:2538-2570: the 3x3 orientation matrix alone (for tests). */
void oc_rodrigues_matrix(const double r[3], double Q[9]);

/* f90:655-706: cmod for every phase.  flag_m=1: grains from rot36, last phase from
 * phasemod_pore = {K, G}.  flag_m=0: phases 1 and 2 from phasemod2 = {K1,G1,K2,G2}. */
void oc_cmod(int nphase, int flag_m, const double *rot36, const double *phasemod, double *cmod);

/* f90:713-805: element stiffness dk by 3x3x3 Simpson quadrature. */
void oc_dk(int nphase, const double *cmod, double *dk);

/* f90:815-1100: periodic-jump right-hand side b (3*ns) and constant C. */
void oc_femat_rhs(int nx, int ny, int nz, const int32_t *pix, const double *dk,
                  const double strain6[6], double *b, double *C);

/* f90:423-438: affine initial guess. */
void oc_affine(int nx, int ny, int nz, const double strain6[6], double *u);

/* f90:1176-1321 / 1585-1884: y = K x, the matrix-free 27-point stencil (the
 * reference's per-neighbour grouping). */
void oc_apply(int nx, int ny, int nz, const int32_t *pix, const double *dk,
              const double *x, double *y);

/* f90:1107-1353: gb = K u + b, utot = C + sum(0.5 u.Ku + b.u). */
void oc_energy(int nx, int ny, int nz, const int32_t *pix, const double *dk, const double *b,
               double C, const double *u, double *gb, double *utot);

/* f90:452: sum(gb^2). */
double oc_sumsq(int64_t n, const double *v);

/* f90:1498-1520 + invert3x3 f90:2046-2147: minv[m*9 + i*3 + j]. */
void oc_invert3x3(const double a[9], double ainv[9]);
void oc_precond(int nx, int ny, int nz, const int32_t *pix, const double *dk, double *minv);

/* f90:1356-2043: one dembx call (kkk = 1).  u, gb, h are in/out (h may be NULL:
 * the reference overwrites it with z before use, f90:1554-1562).  history[ldemb]
 * receives rel_res per iteration.  Returns the OC_* flag; *lstep, *gg set. */
int oc_dembx(int nx, int ny, int nz, const int32_t *pix, const double *dk, double *u,
             double *gb, double *h, double gg0, float tol, int ldemb, int *lstep, double *gg,
             double *history, int *reset_count);

/* The PCG loops of the reference's programs (include/e3x.h e3x_pcg_variant):
 * src/elas3dxtal_pcg.f90 (f90:1356-2043), applications/sec3_2/A_x_y/elas3dxtal_pcg.f90
 * (A_2_0:1045-1558) and applications/sec3_1/cg/n_N/elas3dxtal_cg.f90 (cg:915-1074). */
enum { OC_PCG_REFERENCE = 0, OC_PCG_SWEEP = 1, OC_PCG_PLAIN_CG = 2 };
/* Wall seconds of the last oc_dembx_variant loop (setup excluded). */
double oc_last_loop_seconds(void);
int oc_dembx_variant(int nx, int ny, int nz, const int32_t *pix, const double *dk, double *u,
                     double *gb, double *h, double gg0, float tol, int ldemb, int *lstep, double *gg,
                     double *history, int *reset_count, int variant);

/* f90:2149-2408: volume-averaged stress6 (str*) and engineering strain6 (s*),
 * order xx,yy,zz,xz,yz,xy. */
void oc_stress(int nx, int ny, int nz, const int32_t *pix, const double *cmod,
               const double strain6[6], const double *u, double stress6[6], double strain_avg6[6]);

/* f90:2639-2833: stress_field[c*ns + m] (c = 0..5) and vm[m]. */
void oc_fullfield(int nx, int ny, int nz, const int32_t *pix, const double *cmod,
                  const double strain6[6], const double *u, double *stress_field, double *vm);

/* Summary of program elas3d (f90:326-540) for one microstructure. */
typedef struct {
    double utot0, gg0;          /* f90:449-453 */
    double utot, gg, rel_res;   /* f90:479, 495 */
    int lstep, status, reset_count;
    double C;
    double stress6[6], strain6[6];
} oc_result;

/* Runs femat -> affine -> energy -> dembx -> energy -> stress as the main program does.
 * Outputs u (3*ns) and gb (3*ns) if non-NULL; history[ldemb] if non-NULL. */
int oc_solve(int nx, int ny, int nz, int nphase, int flag_m, const int32_t *pix,
             const double *rod, double c11, double c12, double c44, const double *phasemod,
             const double strain6[6], float tol, int ldemb, double *u_out, double *gb_out,
             double *history, oc_result *res);

/* Hotspots of a von Mises field: applications/sec3_2/A_2_0/elas3dxtal_postprocessing.m
 * :86-215 (solid = label <= nsolid; masked Gaussian smoothing, zero padding, separable
 * passes x, y, z, numerator double and denominator single; regional maxima by flood fill
 * of equal-valued 26-connected components; threshold s/1e6 >= thr_mpa in single;
 * sorted by value descending, ties by index).  Writes the smoothed field to sm if
 * non-NULL and the first npeaks peaks; returns the number of peaks found. */
int64_t oc_hotspots(int nx, int ny, int nz, const int32_t *pix, const double *vm, int nsolid, double sigma,
                    double thr_mpa, int npeaks, int64_t *index, float *value, float *sm);

/* Threads used by the OpenMP loops (0 = library default). */
void oc_set_threads(int n);
int oc_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
