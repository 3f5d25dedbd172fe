/*
 * e3x_oracle.c -- CPU restatement of the elas3dxtal_pcg solve.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker for the HIP library
 * (include/e3x.h).  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product never does.
 *
 * Each function follows one routine of src/elas3dxtal_pcg.f90 (cited).  The
 * reference's OpenMP reductions and `!$omp atomic` scatters have no fixed
 * summation order; here every reduction is a fixed-order chunked sum so the
 * oracle is bitwise repeatable for any thread count.  Built with
 * -ffp-contract=off: plain IEEE double arithmetic in source order.
 */
#include "e3x_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define CHUNK 4096

/* Local node positions of the 8-node brick, from is = (27,3,2,1,26,19,18,17)
 * (f90:639) and the shape-function signs (f90:740-765). */
static const int POS[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                              {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};

/* Neighbour offsets n = 1..27 (f90:326-341). */
static int OFF[27][3];
/* For neighbour n, the (a, b) pairs with POS[b] - POS[a] == OFF[n]: the voxel at
 * m - POS[a] holds node m as local a and the neighbour as local b. */
static int NPAIR[27], PAIR_A[27][8], PAIR_B[27][8];
static int tables_ready = 0;

static void init_tables(void)
{
    if (tables_ready) return;
    const int im8[8] = {0, 1, 1, 1, 0, -1, -1, -1};
    const int jm8[8] = {1, 1, 0, -1, -1, -1, 0, 1};
    for (int n = 0; n < 8; n++) {
        for (int s = 0; s < 3; s++) {
            OFF[n + 8 * s][0] = im8[n];
            OFF[n + 8 * s][1] = jm8[n];
        }
        OFF[n][2] = 0;
        OFF[n + 8][2] = -1;
        OFF[n + 16][2] = 1;
    }
    OFF[24][0] = OFF[24][1] = 0; OFF[24][2] = -1;
    OFF[25][0] = OFF[25][1] = 0; OFF[25][2] = 1;
    OFF[26][0] = OFF[26][1] = OFF[26][2] = 0;
    for (int n = 0; n < 27; n++) {
        NPAIR[n] = 0;
        for (int a = 0; a < 8; a++)
            for (int b = 0; b < 8; b++)
                if (POS[b][0] - POS[a][0] == OFF[n][0] && POS[b][1] - POS[a][1] == OFF[n][1] &&
                    POS[b][2] - POS[a][2] == OFF[n][2]) {
                    PAIR_A[n][NPAIR[n]] = a;
                    PAIR_B[n][NPAIR[n]] = b;
                    NPAIR[n]++;
                }
    }
    tables_ready = 1;
}

static inline int wrapi(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

#define DK(p, a, i, b, j) ((((((size_t)(p)) * 8 + (a)) * 3 + (i)) * 8 + (b)) * 3 + (j))

void oc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int oc_get_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* Fixed-order sum of a chunk-partial array. */
static double sum_partials(const double *p, int64_t n)
{
    double s = 0.0;
    for (int64_t i = 0; i < n; i++) s += p[i];
    return s;
}

/* ---------------------------------------------------------------- f90:326-369 */
void oc_build_ib(int nx, int ny, int nz, int32_t *ib)
{
    init_tables();
    const int64_t ns = (int64_t)nx * ny * nz;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                int64_t m = (int64_t)nx * ny * k + (int64_t)nx * j + i;
                for (int n = 0; n < 27; n++) {
                    int i1 = wrapi(i + OFF[n][0], nx), j1 = wrapi(j + OFF[n][1], ny),
                        k1 = wrapi(k + OFF[n][2], nz);
                    ib[(int64_t)n * ns + m] = (int32_t)((int64_t)nx * ny * k1 + (int64_t)nx * j1 + i1);
                }
            }
}

/* ---------------------------------------------------------------- f90:2410-2468 */
void oc_assig(int64_t ns, int nphase, const int32_t *pix, double *prob)
{
    for (int p = 0; p < nphase; p++) prob[p] = 0.0;
    for (int64_t m = 0; m < ns; m++)
        if (pix[m] >= 1 && pix[m] <= nphase) prob[pix[m] - 1] += 1.0;
    for (int p = 0; p < nphase; p++) prob[p] = prob[p] / (double)ns;
}

/* ---------------------------------------------------------------- f90:2538-2570 */
void oc_rodrigues_matrix(const double r[3], double Q[9])
{
    double rdotr = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
    double term1 = 1.0 - rdotr, term2 = 1.0 + rdotr;
    double om[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) om[i][j] = (i == j) ? term1 : 0.0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) om[i][j] = om[i][j] + (2.0 * r[i]) * r[j];
    om[0][1] = om[0][1] - 2.0 * r[2];
    om[0][2] = om[0][2] + 2.0 * r[1];
    om[1][2] = om[1][2] - 2.0 * r[0];
    om[1][0] = om[1][0] + 2.0 * r[2];
    om[2][0] = om[2][0] - 2.0 * r[1];
    om[2][1] = om[2][1] + 2.0 * r[0];
    if (term2 > 1.0e-12) {
        term2 = 1.0 / term2;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) om[i][j] = om[i][j] * term2;
    } else {
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) om[i][j] = (i == j) ? 1.0 : 0.0;
    }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) Q[i * 3 + j] = om[i][j];
}

/* ---------------------------------------------------------------- f90:2470-2637 */
void oc_rotate_stiffness(int ngrain, const double *rod, double c11, double c12, double c44,
                         double *rot36)
{
    double cc[6][6];
    memset(cc, 0, sizeof cc);
    cc[0][0] = cc[1][1] = cc[2][2] = c11;
    cc[0][1] = cc[0][2] = cc[1][2] = cc[1][0] = cc[2][0] = cc[2][1] = c12;
    cc[3][3] = cc[4][4] = cc[5][5] = c44;
    for (int g = 0; g < ngrain; g++) {
        double q9[9];
        oc_rodrigues_matrix(rod + 3 * g, q9);
#define Q(i, j) q9[((i)-1) * 3 + ((j)-1)]
        double nr[6][6];
        /* rows 1-3 (f90:2578-2597) */
        for (int r = 1; r <= 3; r++) {
            nr[r - 1][0] = Q(r, 1) * Q(r, 1);
            nr[r - 1][1] = Q(r, 2) * Q(r, 2);
            nr[r - 1][2] = Q(r, 3) * Q(r, 3);
            nr[r - 1][3] = Q(r, 2) * Q(r, 3);
            nr[r - 1][4] = Q(r, 3) * Q(r, 1);
            nr[r - 1][5] = Q(r, 1) * Q(r, 2);
        }
        /* row 4 (f90:2600-2605) */
        nr[3][0] = 2.0 * Q(3, 1) * Q(1, 1);
        nr[3][1] = 2.0 * Q(3, 2) * Q(1, 2);
        nr[3][2] = 2.0 * Q(3, 3) * Q(1, 3);
        nr[3][3] = Q(1, 2) * Q(3, 3) + Q(1, 3) * Q(3, 2);
        nr[3][4] = Q(1, 3) * Q(3, 1) + Q(1, 1) * Q(3, 3);
        nr[3][5] = Q(1, 1) * Q(3, 2) + Q(1, 2) * Q(3, 1);
        /* row 5 (f90:2608-2613) */
        nr[4][0] = 2.0 * Q(2, 1) * Q(3, 1);
        nr[4][1] = 2.0 * Q(2, 2) * Q(3, 2);
        nr[4][2] = 2.0 * Q(2, 3) * Q(3, 3);
        nr[4][3] = Q(2, 2) * Q(3, 3) + Q(2, 3) * Q(3, 2);
        nr[4][4] = Q(2, 1) * Q(3, 3) + Q(2, 3) * Q(3, 1);
        nr[4][5] = Q(2, 2) * Q(3, 1) + Q(2, 1) * Q(3, 2);
        /* row 6 (f90:2616-2621) */
        nr[5][0] = 2.0 * Q(1, 1) * Q(2, 1);
        nr[5][1] = 2.0 * Q(1, 2) * Q(2, 2);
        nr[5][2] = 2.0 * Q(1, 3) * Q(2, 3);
        nr[5][3] = Q(1, 2) * Q(2, 3) + Q(1, 3) * Q(2, 2);
        nr[5][4] = Q(1, 3) * Q(2, 1) + Q(1, 1) * Q(2, 3);
        nr[5][5] = Q(1, 1) * Q(2, 2) + Q(1, 2) * Q(2, 1);
#undef Q
        /* rotated = matmul(matmul(transpose(newr), C), newr)  (f90:2627) */
        double t[6][6], rt[6][6];
        for (int i = 0; i < 6; i++)
            for (int j = 0; j < 6; j++) {
                double s = 0.0;
                for (int k = 0; k < 6; k++) s += nr[k][i] * cc[k][j];
                t[i][j] = s;
            }
        for (int i = 0; i < 6; i++)
            for (int j = 0; j < 6; j++) {
                double s = 0.0;
                for (int k = 0; k < 6; k++) s += t[i][k] * nr[k][j];
                rt[i][j] = s;
            }
        for (int i = 0; i < 6; i++)
            for (int l = 0; l < 6; l++) rot36[(size_t)g * 36 + i * 6 + l] = rt[i][l];
    }
}

/* ---------------------------------------------------------------- f90:655-706 */
void oc_cmod(int nphase, int flag_m, const double *rot36, const double *phasemod, double *cmod)
{
    double ck[6][6], cmu[6][6];
    memset(ck, 0, sizeof ck);
    memset(cmu, 0, sizeof cmu);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            ck[i][j] = 1.0;
            cmu[i][j] = (i == j) ? 4.0 / 3.0 : -2.0 / 3.0;
        }
    cmu[3][3] = cmu[4][4] = cmu[5][5] = 1.0;
    memset(cmod, 0, sizeof(double) * 36 * (size_t)nphase);
    if (flag_m == 0) {
        for (int k = 0; k < 2 && k < nphase; k++)
            for (int j = 0; j < 6; j++)
                for (int i = 0; i < 6; i++)
                    cmod[((size_t)k * 6 + i) * 6 + j] =
                        phasemod[2 * k] * ck[i][j] + phasemod[2 * k + 1] * cmu[i][j];
    } else {
        for (int k = 0; k < nphase - 1; k++)
            for (int i = 0; i < 6; i++)
                for (int j = 0; j < 6; j++)
                    cmod[((size_t)k * 6 + i) * 6 + j] = rot36[(size_t)k * 36 + i * 6 + j];
        for (int j = 0; j < 6; j++)
            for (int i = 0; i < 6; i++)
                cmod[((size_t)(nphase - 1) * 6 + i) * 6 + j] =
                    phasemod[0] * ck[i][j] + phasemod[1] * cmu[i][j];
    }
}

/* Strain-displacement matrix at (x,y,z) in the unit brick (f90:740-780):
 * es[r][n8][dof], rows 1=xx 2=yy 3=zz 4=xz 5=yz 6=xy. */
static void build_es(double x, double y, double z, double es[6][8][3])
{
    double dndx[8], dndy[8], dndz[8];
    dndx[0] = -(1.0 - y) * (1.0 - z); dndx[1] = (1.0 - y) * (1.0 - z);
    dndx[2] = y * (1.0 - z);          dndx[3] = -y * (1.0 - z);
    dndx[4] = -(1.0 - y) * z;         dndx[5] = (1.0 - y) * z;
    dndx[6] = y * z;                  dndx[7] = -y * z;
    dndy[0] = -(1.0 - x) * (1.0 - z); dndy[1] = -x * (1.0 - z);
    dndy[2] = x * (1.0 - z);          dndy[3] = (1.0 - x) * (1.0 - z);
    dndy[4] = -(1.0 - x) * z;         dndy[5] = -x * z;
    dndy[6] = x * z;                  dndy[7] = (1.0 - x) * z;
    dndz[0] = -(1.0 - x) * (1.0 - y); dndz[1] = -x * (1.0 - y);
    dndz[2] = -x * y;                 dndz[3] = -(1.0 - x) * y;
    dndz[4] = (1.0 - x) * (1.0 - y);  dndz[5] = x * (1.0 - y);
    dndz[6] = x * y;                  dndz[7] = (1.0 - x) * y;
    memset(es, 0, sizeof(double) * 6 * 8 * 3);
    for (int n8 = 0; n8 < 8; n8++) {
        es[0][n8][0] = dndx[n8];
        es[1][n8][1] = dndy[n8];
        es[2][n8][2] = dndz[n8];
        es[3][n8][0] = dndz[n8];
        es[3][n8][2] = dndx[n8];
        es[4][n8][1] = dndz[n8];
        es[4][n8][2] = dndy[n8];
        es[5][n8][0] = dndy[n8];
        es[5][n8][1] = dndx[n8];
    }
}

/* ---------------------------------------------------------------- f90:713-805 */
void oc_dk(int nphase, const double *cmod, double *dk)
{
    memset(dk, 0, sizeof(double) * 576 * (size_t)nphase);
#pragma omp parallel for schedule(dynamic)
    for (int p = 0; p < nphase; p++) {
        const double *c = cmod + (size_t)p * 36;
        double es[6][8][3];
        for (int k3 = 0; k3 < 3; k3++)
            for (int j3 = 0; j3 < 3; j3++)
                for (int i3 = 0; i3 < 3; i3++) {
                    static const double W[4] = {1.0, 4.0, 16.0, 64.0}; /* 4.0**n (f90:713-721) */
                    double g = W[(i3 == 1) + (j3 == 1) + (k3 == 1)];
                    build_es(i3 / 2.0, j3 / 2.0, k3 / 2.0, es);
                    for (int mm = 0; mm < 3; mm++)
                        for (int nn = 0; nn < 3; nn++)
                            for (int ii = 0; ii < 8; ii++)
                                for (int jj = 0; jj < 8; jj++) {
                                    double tmp[6];
                                    for (int r = 0; r < 6; r++) {
                                        double s = 0.0;
                                        for (int q = 0; q < 6; q++) s += c[r * 6 + q] * es[q][jj][nn];
                                        tmp[r] = s;
                                    }
                                    double sumval = 0.0;
                                    for (int r = 0; r < 6; r++) sumval += es[r][ii][mm] * tmp[r];
                                    dk[DK(p, ii, mm, jj, nn)] += g * sumval / 216.0;
                                }
                }
    }
}

/* One boundary voxel of f90:815-1100: b += K_e delta, and the voxel's constant
 * 1/2 delta^T K_e delta added to the running chunk sum *C (the constant of the voxel first
 * summed on its own: its terms in the reference's order, f90:845-846). */
static void rhs_voxel(int nx, int ny, int nz, int i, int j, int k, const int32_t *pix,
                      const double *dk, const double delta[8][3], double *b, double *C)
{
    const int64_t ns = (int64_t)nx * ny * nz;
    int64_t m = (int64_t)nx * ny * k + (int64_t)nx * j + i;
    int p = pix[m] - 1;
    double cv = 0.0;
    for (int nn = 0; nn < 3; nn++)
        for (int mm = 0; mm < 8; mm++) {
            double sumval = 0.0;
            for (int m3 = 0; m3 < 3; m3++)
                for (int m8 = 0; m8 < 8; m8++) {
                    sumval = sumval + delta[m8][m3] * dk[DK(p, m8, m3, mm, nn)];
                    cv = cv + 0.5 * delta[m8][m3] * dk[DK(p, m8, m3, mm, nn)] * delta[mm][nn];
                }
            int i1 = wrapi(i + POS[mm][0], nx), j1 = wrapi(j + POS[mm][1], ny),
                k1 = wrapi(k + POS[mm][2], nz);
            int64_t ind = (int64_t)nx * ny * k1 + (int64_t)nx * j1 + i1;
            b[nn * ns + ind] += sumval;
        }
    *C = *C + cv;
}

/* Jump of a voxel on the upper x/y/z boundaries (f90:823-1080): local node a gets the
 * sum, in x,y,z order, of E_f = (e_fx, e_fy, e_fz) * n_f over the crossed faces f. */
static void jump_delta(int fx, int fy, int fz, int nx, int ny, int nz, const double e6[6],
                       double delta[8][3])
{
    const double exx = e6[0], eyy = e6[1], ezz = e6[2], exz = e6[3], eyz = e6[4], exy = e6[5];
    double E[3][3] = {{exx * nx, exy * nx, exz * nx},
                      {exy * ny, eyy * ny, eyz * ny},
                      {exz * nz, eyz * nz, ezz * nz}};
    int f[3] = {fx, fy, fz};
    for (int a = 0; a < 8; a++)
        for (int c = 0; c < 3; c++) {
            double s = 0.0;
            int any = 0;
            for (int d = 0; d < 3; d++)
                if (f[d] && POS[a][d]) {
                    s = any ? s + E[d][c] : E[d][c];
                    any = 1;
                }
            delta[a][c] = s;
        }
}

/* ---------------------------------------------------------------- f90:815-1100 */
void oc_femat_rhs(int nx, int ny, int nz, const int32_t *pix, const double *dk,
                  const double strain6[6], double *b, double *C)
{
    init_tables();
    const int64_t ns = (int64_t)nx * ny * nz;
    memset(b, 0, sizeof(double) * 3 * (size_t)ns);
    /* C: the reference sums every term into one running total per OpenMP thread
     * (reduction(+:C), f90:833); over the ~1e5 voxels x 576 terms of a 200^3 grid such a
     * running total drifts by ~1e-12 of C, which the energy (C and b.u cancel to ~1/200 of
     * C) turns into ~2e-10.  Here each voxel's constant is summed on its own, those of a
     * grid line in a running sum, and the lines' sums in a second running sum: within
     * ~1e-16 of the exact total (pinned against an exact sum). */
    double c = 0.0, delta[8][3];
    /* face x = nx (f90:821-854) */
    jump_delta(1, 0, 0, nx, ny, nz, strain6, delta);
    for (int j = 0; j < ny - 1; j++) {
        double cl = 0.0;
        for (int k = 0; k < nz - 1; k++) rhs_voxel(nx, ny, nz, nx - 1, j, k, pix, dk, delta, b, &cl);
        c = c + cl;
    }
    /* face y = ny (f90:856-886) */
    jump_delta(0, 1, 0, nx, ny, nz, strain6, delta);
    for (int i = 0; i < nx - 1; i++) {
        double cl = 0.0;
        for (int k = 0; k < nz - 1; k++) rhs_voxel(nx, ny, nz, i, ny - 1, k, pix, dk, delta, b, &cl);
        c = c + cl;
    }
    /* face z = nz (f90:888-918) */
    jump_delta(0, 0, 1, nx, ny, nz, strain6, delta);
    for (int i = 0; i < nx - 1; i++) {
        double cl = 0.0;
        for (int j = 0; j < ny - 1; j++) rhs_voxel(nx, ny, nz, i, j, nz - 1, pix, dk, delta, b, &cl);
        c = c + cl;
    }
    /* edge x=nx, y=ny (f90:920-959) */
    jump_delta(1, 1, 0, nx, ny, nz, strain6, delta);
    {
        double cl = 0.0;
        for (int k = 0; k < nz - 1; k++) rhs_voxel(nx, ny, nz, nx - 1, ny - 1, k, pix, dk, delta, b, &cl);
        c = c + cl;
    }
    /* edge x=nx, z=nz (f90:961-999) */
    jump_delta(1, 0, 1, nx, ny, nz, strain6, delta);
    {
        double cl = 0.0;
        for (int j = 0; j < ny - 1; j++) rhs_voxel(nx, ny, nz, nx - 1, j, nz - 1, pix, dk, delta, b, &cl);
        c = c + cl;
    }
    /* edge y=ny, z=nz (f90:1001-1039) */
    jump_delta(0, 1, 1, nx, ny, nz, strain6, delta);
    {
        double cl = 0.0;
        for (int i = 0; i < nx - 1; i++) rhs_voxel(nx, ny, nz, i, ny - 1, nz - 1, pix, dk, delta, b, &cl);
        c = c + cl;
    }
    /* corner (f90:1041-1094) */
    jump_delta(1, 1, 1, nx, ny, nz, strain6, delta);
    rhs_voxel(nx, ny, nz, nx - 1, ny - 1, nz - 1, pix, dk, delta, b, &c);
    (void)ns;
    *C = c;
}

/* ---------------------------------------------------------------- f90:423-438 */
void oc_affine(int nx, int ny, int nz, const double e6[6], double *u)
{
    const int64_t ns = (int64_t)nx * ny * nz;
    const double exx = e6[0], eyy = e6[1], ezz = e6[2], exz = e6[3], eyz = e6[4], exy = e6[5];
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                int64_t m = (int64_t)nx * ny * k + (int64_t)nx * j + i;
                double x = (double)i, y = (double)j, z = (double)k;
                u[m] = x * exx + y * exy + z * exz;
                u[ns + m] = x * exy + y * eyy + z * eyz;
                u[2 * ns + m] = x * exz + y * eyz + z * ezz;
            }
}

/* ---------------------------------------- f90:1176-1321 and f90:1585-1884 (same form) */
void oc_apply(int nx, int ny, int nz, const int32_t *pix, const double *dk, const double *x,
              double *y)
{
    init_tables();
    const int64_t ns = (int64_t)nx * ny * nz;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                int64_t m = (int64_t)nx * ny * k + (int64_t)nx * j + i;
                int pv[8];
                int64_t nb[27];
                for (int a = 0; a < 8; a++) {
                    int64_t v = (int64_t)nx * ny * wrapi(k - POS[a][2], nz) +
                                (int64_t)nx * wrapi(j - POS[a][1], ny) + wrapi(i - POS[a][0], nx);
                    pv[a] = pix[v] - 1;
                }
                for (int n = 0; n < 27; n++)
                    nb[n] = (int64_t)nx * ny * wrapi(k + OFF[n][2], nz) +
                            (int64_t)nx * wrapi(j + OFF[n][1], ny) + wrapi(i + OFF[n][0], nx);
                for (int jd = 0; jd < 3; jd++) {
                    double s = 0.0;
                    for (int nd = 0; nd < 3; nd++)
                        for (int n = 0; n < 27; n++) {
                            double c = 0.0;
                            for (int t = 0; t < NPAIR[n]; t++) {
                                int a = PAIR_A[n][t], bb = PAIR_B[n][t];
                                double d = dk[DK(pv[a], a, jd, bb, nd)];
                                c = t ? c + d : d;
                            }
                            s = s + x[nd * ns + nb[n]] * c;
                        }
                    y[jd * ns + m] = s;
                }
            }
}

/* ---------------------------------------------------------------- f90:1107-1353 */
void oc_energy(int nx, int ny, int nz, const int32_t *pix, const double *dk, const double *b,
               double C, const double *u, double *gb, double *utot)
{
    const int64_t ns = (int64_t)nx * ny * nz;
    oc_apply(nx, ny, nz, pix, dk, u, gb);
    int64_t nch = (ns + CHUNK - 1) / CHUNK;
    double *part = (double *)malloc(sizeof(double) * 3 * (size_t)nch);
    for (int m3 = 0; m3 < 3; m3++) {
#pragma omp parallel for schedule(static)
        for (int64_t c = 0; c < nch; c++) {
            double s = 0.0;
            int64_t m1 = (c + 1) * CHUNK < ns ? (c + 1) * CHUNK : ns;
            for (int64_t m = c * CHUNK; m < m1; m++) {
                int64_t q = m3 * ns + m;
                s = s + 0.5 * u[q] * gb[q] + b[q] * u[q];
                gb[q] = gb[q] + b[q];
            }
            part[m3 * nch + c] = s;
        }
    }
    *utot = C + sum_partials(part, 3 * nch);
    free(part);
}

double oc_sumsq(int64_t n, const double *v)
{
    int64_t nch = (n + CHUNK - 1) / CHUNK;
    double *part = (double *)malloc(sizeof(double) * (size_t)(nch ? nch : 1));
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < nch; c++) {
        double s = 0.0;
        int64_t m1 = (c + 1) * CHUNK < n ? (c + 1) * CHUNK : n;
        for (int64_t m = c * CHUNK; m < m1; m++) s += v[m] * v[m];
        part[c] = s;
    }
    double r = sum_partials(part, nch);
    free(part);
    return r;
}

/* ---------------------------------------------------------------- f90:2046-2147 */
void oc_invert3x3(const double a9[9], double ainv9[9])
{
#define A(i, j) a9[((i)-1) * 3 + ((j)-1)]
#define AI(i, j) ainv9[((i)-1) * 3 + ((j)-1)]
    double c11 = A(2, 2) * A(3, 3) - A(2, 3) * A(3, 2);
    double c21 = -A(2, 1) * A(3, 3) + A(2, 3) * A(3, 1);
    double c31 = A(2, 1) * A(3, 2) - A(2, 2) * A(3, 1);
    double det = A(1, 1) * c11 + A(1, 2) * c21 + A(1, 3) * c31;
    if (fabs(det) < 1e-14) {
        for (int t = 0; t < 9; t++) ainv9[t] = 0.0;
        return;
    }
    double inv_det = 1.0 / det;
    double c12 = -A(1, 2) * A(3, 3) + A(1, 3) * A(3, 2);
    double c22 = A(1, 1) * A(3, 3) - A(1, 3) * A(3, 1);
    double c32 = -A(1, 1) * A(3, 2) + A(1, 2) * A(3, 1);
    double c13 = A(1, 2) * A(2, 3) - A(1, 3) * A(2, 2);
    double c23 = -A(1, 1) * A(2, 3) + A(1, 3) * A(2, 1);
    double c33 = A(1, 1) * A(2, 2) - A(1, 2) * A(2, 1);
    AI(1, 1) = c11 * inv_det; AI(2, 1) = c21 * inv_det; AI(3, 1) = c31 * inv_det;
    AI(1, 2) = c12 * inv_det; AI(2, 2) = c22 * inv_det; AI(3, 2) = c32 * inv_det;
    AI(1, 3) = c13 * inv_det; AI(2, 3) = c23 * inv_det; AI(3, 3) = c33 * inv_det;
#undef A
#undef AI
}

/* ---------------------------------------------------------------- f90:1498-1520 */
void oc_precond(int nx, int ny, int nz, const int32_t *pix, const double *dk, double *minv)
{
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                int64_t m = (int64_t)nx * ny * k + (int64_t)nx * j + i;
                int pv[8];
                for (int a = 0; a < 8; a++) {
                    int64_t v = (int64_t)nx * ny * wrapi(k - POS[a][2], nz) +
                                (int64_t)nx * wrapi(j - POS[a][1], ny) + wrapi(i - POS[a][0], nx);
                    pv[a] = pix[v] - 1;
                }
                double mb[9];
                for (int ii = 0; ii < 3; ii++)
                    for (int jj = 0; jj < 3; jj++) {
                        double s = dk[DK(pv[0], 0, ii, 0, jj)];
                        for (int a = 1; a < 8; a++) s = s + dk[DK(pv[a], a, ii, a, jj)];
                        mb[ii * 3 + jj] = s;
                    }
                oc_invert3x3(mb, minv + (size_t)m * 9);
            }
}

/* z = M^-1 g (f90:1527-1535, 1927-1935) */
static void apply_minv(int64_t ns, const double *minv, const double *g, double *z)
{
#pragma omp parallel for schedule(static)
    for (int64_t m = 0; m < ns; m++) {
        const double *mi = minv + (size_t)m * 9;
        double g1 = g[m], g2 = g[ns + m], g3 = g[2 * ns + m];
        for (int i = 0; i < 3; i++) z[i * ns + m] = mi[i * 3 + 0] * g1 + mi[i * 3 + 1] * g2 + mi[i * 3 + 2] * g3;
    }
}

/* Fixed-order sum_{c,m} a*b over 3 components (f90:1541-1548, 1889-1896, 1939-1946). */
static double dot3(int64_t ns, const double *a, const double *b, double *part)
{
    int64_t nch = (ns + CHUNK - 1) / CHUNK;
    for (int c3 = 0; c3 < 3; c3++) {
#pragma omp parallel for schedule(static)
        for (int64_t c = 0; c < nch; c++) {
            double s = 0.0;
            int64_t m1 = (c + 1) * CHUNK < ns ? (c + 1) * CHUNK : ns;
            for (int64_t m = c * CHUNK; m < m1; m++) s = s + a[c3 * ns + m] * b[c3 * ns + m];
            part[c3 * nch + c] = s;
        }
    }
    return sum_partials(part, 3 * nch);
}

/* ---------------------------------------------------------------- f90:1356-2043 */
int oc_dembx(int nx, int ny, int nz, const int32_t *pix, const double *dk, double *u,
             double *gb, double *h_io, double gg0, float tol, int ldemb, int *lstep, double *gg_out,
             double *history, int *reset_count_out)
{
    return oc_dembx_variant(nx, ny, nz, pix, dk, u, gb, h_io, gg0, tol, ldemb, lstep, gg_out, history,
                            reset_count_out, OC_PCG_REFERENCE);
}

/* The three loops differ only where marked: the preconditioner (plain CG: none, z = gb,
 * h = gb at entry, cg:949), the divergence reset (src only) and the order of the stop
 * test and the direction update (src: update first, f90:1996/2007; sweep and plain CG:
 * test first, A_2_0:1518-1526, cg:1050-1056). */
/* Wall seconds of the iteration loop (f90:1574-2012) of the last oc_dembx_variant call,
 * setup (preconditioner, z, h) excluded: the CPU baseline's timing scope. */
static double g_loop_seconds = 0.0;
double oc_last_loop_seconds(void) { return g_loop_seconds; }

int oc_dembx_variant(int nx, int ny, int nz, const int32_t *pix, const double *dk, double *u,
                     double *gb, double *h_io, double gg0, float tol, int ldemb, int *lstep,
                     double *gg_out, double *history, int *reset_count_out, int variant)
{
    const int plain = variant == OC_PCG_PLAIN_CG, test_first = variant != OC_PCG_REFERENCE;
    const int64_t ns = (int64_t)nx * ny * nz;
    const double tol_d = (double)tol;
    int64_t nch = (ns + CHUNK - 1) / CHUNK;
    double *part = (double *)malloc(sizeof(double) * 3 * (size_t)nch);
    double *minv = (double *)malloc(sizeof(double) * 9 * (size_t)ns);
    double *z = (double *)malloc(sizeof(double) * 3 * (size_t)ns);
    double *ah = (double *)malloc(sizeof(double) * 3 * (size_t)ns);
    double *h = h_io ? h_io : (double *)malloc(sizeof(double) * 3 * (size_t)ns);

    if (!plain) oc_precond(nx, ny, nz, pix, dk, minv);  /* kkk == 1 (f90:1498-1520) */
    if (plain) memcpy(z, gb, sizeof(double) * 3 * (size_t)ns);  /* h = gb (cg:949) */
    else apply_minv(ns, minv, gb, z);                 /* f90:1527-1535 */
    double gg = dot3(ns, z, gb, part);                /* f90:1541-1548 */
    memcpy(h, z, sizeof(double) * 3 * (size_t)ns);    /* f90:1554-1562 */
    double rel_res_prev = sqrt(gg / gg0), rel_res_min = rel_res_prev, rel_res = rel_res_prev;
    int reset_count = 0, L = 0, status = OC_MAXITER;
    (void)rel_res_min;

    const double t_loop = omp_get_wtime();
    for (int ijk = 1; ijk <= ldemb; ijk++) {
        L = L + 1;
        oc_apply(nx, ny, nz, pix, dk, h, ah);                 /* f90:1578-1883 */
        double hah = dot3(ns, h, ah, part);                   /* f90:1889-1896 */
        if (fabs(hah) < 1e-14) { status = OC_HAH_BREAK; break; } /* f90:1899-1902 */
        double lambda = gg / hah;                             /* f90:1907 */
#pragma omp parallel for schedule(static)
        for (int64_t q = 0; q < 3 * ns; q++) {                /* f90:1914-1921 */
            u[q] = u[q] - lambda * h[q];
            gb[q] = gb[q] - lambda * ah[q];
        }
        if (plain) memcpy(z, gb, sizeof(double) * 3 * (size_t)ns);
        else apply_minv(ns, minv, gb, z);                     /* f90:1927-1935 */
        double gglast = gg;
        gg = dot3(ns, z, gb, part);                           /* f90:1938-1946; cg:1040-1046 */
        rel_res = sqrt(gg / gg0);                             /* f90:1949 */
        if (history) history[L - 1] = rel_res;
        if (test_first && rel_res < tol_d) { status = OC_CONVERGED; break; } /* A_2_0:1518, cg:1050 */
        if (!test_first && rel_res > 10.0 * rel_res_prev && L > 1) { /* f90:1961-1981 */
            memcpy(h, z, sizeof(double) * 3 * (size_t)ns);
            rel_res_min = rel_res;
            reset_count = reset_count + 1;
        } else {                                              /* f90:1983-2001 */
            double gamma = gg / gglast;
#pragma omp parallel for schedule(static)
            for (int64_t q = 0; q < 3 * ns; q++) h[q] = z[q] + gamma * h[q];
            if (rel_res < rel_res_min) rel_res_min = rel_res;
        }
        rel_res_prev = rel_res;                               /* f90:2004 */
        if (rel_res < tol_d) { status = OC_CONVERGED; break; } /* f90:2007 (reached by src only) */
    }
    g_loop_seconds = omp_get_wtime() - t_loop;
    *lstep = L;
    *gg_out = gg;
    if (reset_count_out) *reset_count_out = reset_count;
    if (!h_io) free(h);
    free(part); free(minv); free(z); free(ah);
    return status;
}

/* Gather the 8 node displacements of voxel (i,j,k) with the periodic jumps
 * (f90:2260-2323, 2732-2796). */
static void gather_uu(int nx, int ny, int nz, int i, int j, int k, const double e6[6],
                      const double *u, double uu[8][3])
{
    const int64_t ns = (int64_t)nx * ny * nz;
    const double exx = e6[0], eyy = e6[1], ezz = e6[2], exz = e6[3], eyz = e6[4], exy = e6[5];
    for (int a = 0; a < 8; a++) {
        int64_t q = (int64_t)nx * ny * wrapi(k + POS[a][2], nz) + (int64_t)nx * wrapi(j + POS[a][1], ny) +
                    wrapi(i + POS[a][0], nx);
        for (int c = 0; c < 3; c++) uu[a][c] = u[c * ns + q];
    }
    if (i == nx - 1) {
        const int L[4] = {1, 2, 5, 6};
        for (int t = 0; t < 4; t++) uu[L[t]][0] = uu[L[t]][0] + exx * (double)nx;
        for (int t = 0; t < 4; t++) uu[L[t]][1] = uu[L[t]][1] + exy * (double)nx;
        for (int t = 0; t < 4; t++) uu[L[t]][2] = uu[L[t]][2] + exz * (double)nx;
    }
    if (j == ny - 1) {
        const int L[4] = {2, 3, 6, 7};
        for (int t = 0; t < 4; t++) uu[L[t]][0] = uu[L[t]][0] + exy * (double)ny;
        for (int t = 0; t < 4; t++) uu[L[t]][1] = uu[L[t]][1] + eyy * (double)ny;
        for (int t = 0; t < 4; t++) uu[L[t]][2] = uu[L[t]][2] + eyz * (double)ny;
    }
    if (k == nz - 1) {
        const int L[4] = {4, 5, 6, 7};
        for (int t = 0; t < 4; t++) uu[L[t]][0] = uu[L[t]][0] + exz * (double)nz;
        for (int t = 0; t < 4; t++) uu[L[t]][1] = uu[L[t]][1] + eyz * (double)nz;
        for (int t = 0; t < 4; t++) uu[L[t]][2] = uu[L[t]][2] + ezz * (double)nz;
    }
}

/* ---------------------------------------------------------------- f90:2149-2408 */
void oc_stress(int nx, int ny, int nz, const int32_t *pix, const double *cmod,
               const double strain6[6], const double *u, double stress6[6], double strain_avg6[6])
{
    const int64_t ns = (int64_t)nx * ny * nz;
    double es[6][8][3];
    build_es(0.5, 0.5, 0.5, es); /* +-0.25 centre derivatives (f90:2224-2238) */
    int64_t nplane = (int64_t)ny * nz;
    double *part = (double *)calloc((size_t)nplane * 12, sizeof(double));
#pragma omp parallel for schedule(static)
    for (int64_t pl = 0; pl < nplane; pl++) {
        int j = (int)(pl % ny), k = (int)(pl / ny);
        double acc[12] = {0};
        for (int i = 0; i < nx; i++) {
            int64_t m = (int64_t)nx * ny * k + (int64_t)nx * j + i;
            double uu[8][3];
            gather_uu(nx, ny, nz, i, j, k, strain6, u, uu);
            const double *c = cmod + (size_t)(pix[m] - 1) * 36;
            double t[12] = {0};
            for (int n3 = 0; n3 < 3; n3++)
                for (int n8 = 0; n8 < 8; n8++) {
                    for (int r = 0; r < 6; r++) t[6 + r] = t[6 + r] + es[r][n8][n3] * uu[n8][n3];
                    for (int n = 0; n < 6; n++)
                        for (int r = 0; r < 6; r++)
                            t[r] = t[r] + c[r * 6 + n] * es[n][n8][n3] * uu[n8][n3];
                }
            for (int r = 0; r < 12; r++) acc[r] = acc[r] + t[r];
        }
        for (int r = 0; r < 12; r++) part[pl * 12 + r] = acc[r];
    }
    double tot[12] = {0};
    for (int64_t pl = 0; pl < nplane; pl++)
        for (int r = 0; r < 12; r++) tot[r] += part[pl * 12 + r];
    for (int r = 0; r < 6; r++) {
        stress6[r] = tot[r] / (double)ns;
        strain_avg6[r] = tot[6 + r] / (double)ns;
    }
    free(part);
}

/* ---------------------------------------------------------------- f90:2639-2833 */
void oc_fullfield(int nx, int ny, int nz, const int32_t *pix, const double *cmod,
                  const double strain6[6], const double *u, double *sf, double *vm)
{
    const int64_t ns = (int64_t)nx * ny * nz;
    double es[6][8][3];
    build_es(0.5, 0.5, 0.5, es);
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                int64_t m = (int64_t)nx * ny * k + (int64_t)nx * j + i;
                double uu[8][3], eps[6] = {0};
                gather_uu(nx, ny, nz, i, j, k, strain6, u, uu);
                for (int n3 = 0; n3 < 3; n3++)
                    for (int n8 = 0; n8 < 8; n8++)
                        for (int r = 0; r < 6; r++) eps[r] = eps[r] + es[r][n8][n3] * uu[n8][n3];
                const double *c = cmod + (size_t)(pix[m] - 1) * 36;
                for (int r = 0; r < 6; r++) {
                    double s = 0.0;
                    for (int q = 0; q < 6; q++) s += c[r * 6 + q] * eps[q];
                    sf[r * ns + m] = s;
                }
            }
#pragma omp parallel for schedule(static)
    for (int64_t m = 0; m < ns; m++) {
        double s1 = sf[m], s2 = sf[ns + m], s3 = sf[2 * ns + m], s4 = sf[3 * ns + m],
               s5 = sf[4 * ns + m], s6 = sf[5 * ns + m];
        vm[m] = sqrt(0.5 * ((s1 - s2) * (s1 - s2) + (s2 - s3) * (s2 - s3) + (s3 - s1) * (s3 - s1) +
                            6.0 * (s4 * s4 + s5 * s5 + s6 * s6)));
    }
}

/* ---------------------------------------------------------------- f90:326-540 */
int oc_solve(int nx, int ny, int nz, int nphase, int flag_m, const int32_t *pix,
             const double *rod, double c11, double c12, double c44, const double *phasemod,
             const double strain6[6], float tol, int ldemb, double *u_out, double *gb_out,
             double *history, oc_result *res)
{
    const int64_t ns = (int64_t)nx * ny * nz;
    for (int64_t m = 0; m < ns; m++)
        if (pix[m] < 1 || pix[m] > nphase) return -1; /* f90:2986-2989 */
    double *rot = (double *)calloc((size_t)36 * (nphase > 1 ? nphase - 1 : 1), sizeof(double));
    double *cmod = (double *)malloc(sizeof(double) * 36 * (size_t)nphase);
    double *dk = (double *)malloc(sizeof(double) * 576 * (size_t)nphase);
    double *b = (double *)malloc(sizeof(double) * 3 * (size_t)ns);
    double *u = u_out ? u_out : (double *)malloc(sizeof(double) * 3 * (size_t)ns);
    double *gb = gb_out ? gb_out : (double *)malloc(sizeof(double) * 3 * (size_t)ns);
    double *h = (double *)malloc(sizeof(double) * 3 * (size_t)ns);
    if (flag_m == 1) oc_rotate_stiffness(nphase - 1, rod, c11, c12, c44, rot);
    oc_cmod(nphase, flag_m, rot, phasemod, cmod);
    oc_dk(nphase, cmod, dk);
    double C;
    oc_femat_rhs(nx, ny, nz, pix, dk, strain6, b, &C);
    oc_affine(nx, ny, nz, strain6, u);
    double utot;
    oc_energy(nx, ny, nz, pix, dk, b, C, u, gb, &utot);
    double gg = oc_sumsq(3 * ns, gb), gg0 = gg;
    res->utot0 = utot;
    res->gg0 = gg0;
    res->C = C;
    int L = 0, rc = 0;
    res->status = oc_dembx(nx, ny, nz, pix, dk, u, gb, h, gg0, tol, ldemb, &L, &gg, history, &rc);
    oc_energy(nx, ny, nz, pix, dk, b, C, u, gb, &utot);
    res->utot = utot;
    res->gg = gg;
    res->rel_res = sqrt(gg / gg0);
    res->lstep = L;
    res->reset_count = rc;
    oc_stress(nx, ny, nz, pix, cmod, strain6, u, res->stress6, res->strain6);
    if (!u_out) free(u);
    if (!gb_out) free(gb);
    free(rot); free(cmod); free(dk); free(b); free(h);
    return 0;
}

/* ---------------------------------------------------------------- postprocessing.m:86-243 */
typedef struct {
    float v;
    int64_t m;
} oc_peak;

static int oc_peak_cmp(const void *a, const void *b)
{
    const oc_peak *x = (const oc_peak *)a, *y = (const oc_peak *)b;
    if (x->v != y->v) return x->v > y->v ? -1 : 1;  /* sort(..., 'descend') (.m:215) */
    return x->m < y->m ? -1 : (x->m > y->m ? 1 : 0); /* stable: find() order */
}

/* One separable zero-padded pass of a (2r+1) kernel along axis `ax` (imgaussfilt3 with
 * 'Padding', 0 filters each dimension in turn). */
static void conv_axis_d(int nx, int ny, int nz, int ax, int r, const double *w, const double *in, double *out)
{
    const int n[3] = {nx, ny, nz};
    const int64_t stride[3] = {1, nx, (int64_t)nx * ny};
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                const int c[3] = {i, j, k};
                const int64_t m = i + (int64_t)nx * (j + (int64_t)ny * k);
                double acc = 0.0;
                for (int d = -r; d <= r; d++) {
                    const int p = c[ax] + d;
                    if (p < 0 || p >= n[ax]) continue;
                    acc += w[d < 0 ? -d : d] * in[m + d * stride[ax]];
                }
                out[m] = acc;
            }
}

static void conv_axis_f(int nx, int ny, int nz, int ax, int r, const float *w, const float *in, float *out)
{
    const int n[3] = {nx, ny, nz};
    const int64_t stride[3] = {1, nx, (int64_t)nx * ny};
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                const int c[3] = {i, j, k};
                const int64_t m = i + (int64_t)nx * (j + (int64_t)ny * k);
                float acc = 0.0f;
                for (int d = -r; d <= r; d++) {
                    const int p = c[ax] + d;
                    if (p < 0 || p >= n[ax]) continue;
                    acc += w[d < 0 ? -d : d] * in[m + d * stride[ax]];
                }
                out[m] = acc;
            }
}

int64_t oc_hotspots(int nx, int ny, int nz, const int32_t *pix, const double *vm, int nsolid, double sigma,
                    double thr_mpa, int npeaks, int64_t *index, float *value, float *sm_out)
{
    const int64_t ns = (int64_t)nx * ny * nz;
    const int r = (int)ceil(2.0 * sigma);
    double w[16], tot = 0.0;
    float wf[16];
    for (int d = -r; d <= r; d++) tot += exp(-(double)d * d / (2.0 * sigma * sigma));
    for (int d = 0; d <= r; d++) {
        w[d] = exp(-(double)d * d / (2.0 * sigma * sigma)) / tot;
        wf[d] = (float)w[d];
    }
    double *a = (double *)malloc(sizeof(double) * (size_t)ns), *b = (double *)malloc(sizeof(double) * (size_t)ns);
    float *fa = (float *)malloc(sizeof(float) * (size_t)ns), *fb = (float *)malloc(sizeof(float) * (size_t)ns);
    float *sm = (float *)malloc(sizeof(float) * (size_t)ns);
    for (int64_t m = 0; m < ns; m++) {                  /* .m:86-88, 92 */
        const int sol = pix[m] <= nsolid;
        a[m] = sol ? vm[m] : 0.0;
        fa[m] = sol ? 1.0f : 0.0f;
    }
    conv_axis_d(nx, ny, nz, 0, r, w, a, b);              /* .m:91 */
    conv_axis_d(nx, ny, nz, 1, r, w, b, a);
    conv_axis_d(nx, ny, nz, 2, r, w, a, b);
    conv_axis_f(nx, ny, nz, 0, r, wf, fa, fb);           /* .m:92 */
    conv_axis_f(nx, ny, nz, 1, r, wf, fb, fa);
    conv_axis_f(nx, ny, nz, 2, r, wf, fa, fb);
    for (int64_t m = 0; m < ns; m++) {
        float den = fb[m];
        if (den < 1e-6f) den = 1.0f;                     /* .m:93 */
        sm[m] = pix[m] <= nsolid ? (float)b[m] / den : 0.0f; /* .m:96-97 */
    }
    /* imregionalmax (.m:206): flood-fill each equal-valued component */
    uint8_t *seen = (uint8_t *)calloc((size_t)ns, 1);
    int64_t *stack = (int64_t *)malloc(sizeof(int64_t) * (size_t)ns);
    int64_t *comp = (int64_t *)malloc(sizeof(int64_t) * (size_t)ns);
    oc_peak *pk = (oc_peak *)malloc(sizeof(oc_peak) * (size_t)ns);
    int64_t npk = 0;
    for (int64_t s0 = 0; s0 < ns; s0++) {
        if (seen[s0]) continue;
        const float v = sm[s0];
        int64_t top = 0, nc = 0;
        int is_max = 1;
        stack[top++] = s0;
        seen[s0] = 1;
        while (top > 0) {
            const int64_t m = stack[--top];
            comp[nc++] = m;
            const int i = (int)(m % nx), j = (int)((m / nx) % ny), k = (int)(m / ((int64_t)nx * ny));
            for (int dz = -1; dz <= 1; dz++)
                for (int dy = -1; dy <= 1; dy++)
                    for (int dx = -1; dx <= 1; dx++) {
                        const int ii = i + dx, jj = j + dy, kk = k + dz;
                        if (ii < 0 || ii >= nx || jj < 0 || jj >= ny || kk < 0 || kk >= nz) continue;
                        const int64_t q = ii + (int64_t)nx * (jj + (int64_t)ny * kk);
                        if (sm[q] > v) is_max = 0;
                        if (sm[q] == v && !seen[q]) {
                            seen[q] = 1;
                            stack[top++] = q;
                        }
                    }
        }
        if (!is_max || !(v / 1e6f >= (float)thr_mpa)) continue; /* .m:207-208 */
        for (int64_t c = 0; c < nc; c++) {
            pk[npk].v = v;
            pk[npk].m = comp[c];
            npk++;
        }
    }
    qsort(pk, (size_t)npk, sizeof(oc_peak), oc_peak_cmp);
    for (int p = 0; p < npeaks; p++) {
        index[p] = p < npk ? pk[p].m : -1;
        value[p] = p < npk ? pk[p].v : 0.0f;
    }
    if (sm_out) memcpy(sm_out, sm, sizeof(float) * (size_t)ns);
    free(a); free(b); free(fa); free(fb); free(sm); free(seen); free(stack); free(comp); free(pk);
    return npk;
}
