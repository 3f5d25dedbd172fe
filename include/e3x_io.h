/*
 * e3x_io.h -- host-side input/output of the elas3dxtal_pcg solver (libe3x_io.so).
 *
 *  - HDF5 input and output in the reference's schema:
 *      input_structure_poly.h5 {/pix int32[ns], /orientation f64[3*ngrain]}
 *        read by ppixel_hdf5 (src/elas3dxtal_pcg.f90:2879-2991), written by the
 *        reference's MATLAB generator (src/microstructure_gen.m:867-889);
 *      fullfield_poly.h5 {/pix int32[ns], /vm f64[ns], /stress_tensor f64 (ns,6)}
 *        written by stress_fullfield_OpenMP (f90:2840-2874).
 *    HDF5 1.10 is loaded at run time (dlopen of libhdf5.so.103; E3X_EIO if absent).
 *  - A seeded periodic microstructure generator for synthetic inputs (periodic
 *    Voronoi grains, uniform SO(3) orientations, pores).  It replaces the
 *    reference's MATLAB generator (src/microstructure_gen.m), which is unseeded and
 *    non-periodic, for producing inputs in the same schema.
 *
 * Status codes are those of e3x.h (E3X_OK, E3X_EINVAL, E3X_EPHASE, E3X_EIO).
 */
#ifndef E3X_IO_H
#define E3X_IO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Message of the last failing call of this library on this thread. */
const char *e3x_io_last_error(void);

/* --------------------------------------------------------------- HDF5 (reference schema) */

/* Sizes of an input file: number of /pix elements and /orientation elements, and
 * the /pix "dimensions" attribute (Nx, Ny, Nz) if present (else zeros). */
int e3x_h5_probe_input(const char *path, int64_t *npix, int64_t *norient, int32_t dims[3]);

/*
 * ppixel_hdf5 (f90:2879-2991): read /pix into pix[ns] and /orientation into
 * orientation[3*(nphase-1)].  Errors as the reference's `stop`s, returned instead:
 * E3X_EIO if the file cannot be opened (f90:2924), E3X_EINVAL if the /pix element
 * count differs from ns (f90:2939-2942) or the /orientation count from
 * 3*(nphase-1) (f90:2966-2969), E3X_EPHASE if a label is outside [1, nphase]
 * (f90:2986-2989).
 */
int e3x_h5_read_input(const char *path, int64_t ns, int nphase, int32_t *pix, double *orientation);

/* Write an input file in the MATLAB generator's layout (microstructure_gen.m:867-889):
 * /pix int32 dims {1, ns} with attribute dimensions = {nx, ny, nz}; /orientation f64
 * dims {1, 3*ngrain} with attribute dimensions = {3*ngrain, 1}. */
int e3x_h5_write_input(const char *path, int nx, int ny, int nz, const int32_t *pix, int ngrain,
                       const double *orientation);

/* stress_fullfield_OpenMP's HDF5 output (f90:2840-2874): /pix int32[ns], /vm f64[ns],
 * /stress_tensor f64 with C dims {6, ns} (Fortran stress_field(ns,6), column-major). */
int e3x_h5_write_fullfield(const char *path, int64_t ns, const int32_t *pix, const double *vm,
                           const double *stress);

/* Read the three datasets of a full-field file back (any pointer may be NULL). */
int e3x_h5_read_fullfield(const char *path, int64_t ns, int32_t *pix, double *vm, double *stress);

/* ------------------------------------------------------------- serving requests */

struct e3x_request;  /* include/e3x_serve.h */

/*
 * A serving request (e3x_serve.h) from files; no counterpart in the reference, whose runs
 * take their parameters from the source.  The microstructure comes from an input file in
 * the reference's schema (as e3x_h5_read_input: the grid from the /pix "dimensions"
 * attribute or the "grid" control, nphase = number of /orientation triples + 1) and the
 * controls from one flat JSON object, every key optional, each default the reference's:
 *   "grid": [nx, ny, nz]            "c11", "c12", "c44": Pa (246.5e9, 147.3e9, 124.7e9)
 *   "zener": A (c11 = c12 + 2 c44 / A; not with "c11")
 *   "strain6": [exx, eyy, ezz, exz, eyz, exy] tensor components (float32(1e-3), 0, ...)
 *   "phasemod": [K, G] of the pore phase (or [K1, G1, K2, G2] for flag_m 0), "flag_m": 1
 *   "precision": "fp64" | "mixed", "replace_every": 50, "tol": float32(1e-8),
 *   "ldemb": 10000, "fullfield": path of the full-field response (fullfield_poly.h5)
 * Unknown keys are errors (E3X_EINVAL).  Fills *rq (its stress_field and vm NULL) and
 * returns the arrays it points to in *pix and *orientation (free them with
 * e3x_request_free once the request is answered), and the "fullfield" path in
 * fullfield_path[path_cap] ("" when absent): the caller then sets rq->stress_field and
 * rq->vm to buffers of 6*ns and ns doubles and, with the response, writes the file with
 * e3x_h5_write_fullfield.
 */
int e3x_request_from_files(const char *h5_path, const char *json, struct e3x_request *rq, int32_t **pix,
                           double **orientation, char *fullfield_path, int path_cap);
void e3x_request_free(int32_t *pix, double *orientation);

/* ------------------------------------------------------------------------ generator */

enum e3x_pore_mode {
    E3X_PORES_NONE = 0,
    E3X_PORES_VOXELS = 1,   /* isolated random voxels until `pore_fraction` of the grid    */
    E3X_PORES_SPHERES = 2   /* lognormal-radius spheres (median, sigma of ln r) until
                               `pore_fraction`                                             */
};

typedef struct {
    int nx, ny, nz;           /* grid (voxels = periodic nodes per axis)                  */
    int ngrain;               /* Voronoi grains; the pore phase is ngrain + 1             */
    uint64_t seed;            /* std::mt19937_64 seed                                     */
    int pore_mode;            /* e3x_pore_mode                                            */
    double pore_fraction;     /* target pore volume fraction of that mode                 */
    double pore_median;       /* lognormal sphere radius median, voxels                   */
    double pore_sigma;        /* sigma of ln(radius)                                      */
    double lof_fraction;      /* extra volume fraction of lack-of-fusion ellipsoids       */
    double lof_axes[3];       /* their semi-axes (voxels) along x', y', z before the yaw  */
} e3x_gen_params;

/*
 * Generate one microstructure: pix[ns] (1..ngrain grains as periodic Voronoi cells of
 * seeded uniform centres, minimum-image distance, ties to the lower grain id; the pore
 * phase ngrain+1), rodrigues[3*ngrain] uniform on SO(3) (Shoemake quaternion, Rodrigues
 * = q_xyz/q_w as microstructure_gen.m:783-801), euler[3*ngrain] Bunge (phi1, Phi, phi2)
 * of the same orientations (may be NULL), *pore_out = achieved pore fraction (may be NULL).
 * Deterministic for a given parameter set.
 */
int e3x_generate(const e3x_gen_params *p, int32_t *pix, double *rodrigues, double *euler,
                 double *pore_out);

/* Bunge (phi1, Phi, phi2) in radians -> Rodrigues vectors r such that the reference's
 * orientation matrix Q(r) (f90:2538-2570) equals the passive Bunge matrix
 * g = Rz(phi2) Rx(Phi) Rz(phi1); and the inverse map. */
int e3x_rodrigues_from_bunge(int n, const double *euler, double *rodrigues);
int e3x_bunge_from_rodrigues(int n, const double *rodrigues, double *euler);

#ifdef __cplusplus
}
#endif
#endif /* E3X_IO_H */
