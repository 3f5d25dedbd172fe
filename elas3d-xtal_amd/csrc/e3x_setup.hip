// gfx950 set-up kernels of the elas3dxtal_pcg solve: femat (per-phase tables), the
// periodic-jump right-hand side and constant, the affine initial guess and the
// block-Jacobi preconditioner.
#include "e3x_device.hpp"

#include <algorithm>

namespace e3x {
namespace {

// ---------------------------------------------------------------------------
// femat: one workgroup per phase (rotatestiffnessby f90:2470-2637, cmod f90:655-706,
// dk by 3x3x3 Simpson f90:713-805) plus the homogeneous 27-point stencil of every
// phase: block S[nb][j][n] = sum over the (voxel, local a, local b) pairs that couple a
// node to its neighbour nb of dk[p][a][j][b][n], stored as its symmetric half (126).
__device__ double shape_deriv(int a, int d, double x, double y, double z)
{
    const double fx = PX(a) ? x : 1.0 - x, fy = PY(a) ? y : 1.0 - y, fz = PZ(a) ? z : 1.0 - z;
    if (d == 0) { double t = fy * fz; return PX(a) ? t : -t; }
    if (d == 1) { double t = fx * fz; return PY(a) ? t : -t; }
    double t = fx * fy;
    return PZ(a) ? t : -t;
}

// es(r, a, i) of f90:769-780; rows xx, yy, zz, xz, yz, xy.
__device__ double es_entry(int r, int a, int i, double x, double y, double z)
{
    switch (r) {
    case 0: return i == 0 ? shape_deriv(a, 0, x, y, z) : 0.0;
    case 1: return i == 1 ? shape_deriv(a, 1, x, y, z) : 0.0;
    case 2: return i == 2 ? shape_deriv(a, 2, x, y, z) : 0.0;
    case 3: return i == 0 ? shape_deriv(a, 2, x, y, z) : (i == 2 ? shape_deriv(a, 0, x, y, z) : 0.0);
    case 4: return i == 1 ? shape_deriv(a, 2, x, y, z) : (i == 2 ? shape_deriv(a, 1, x, y, z) : 0.0);
    default: return i == 0 ? shape_deriv(a, 1, x, y, z) : (i == 1 ? shape_deriv(a, 0, x, y, z) : 0.0);
    }
}

__global__ __launch_bounds__(kBlock) void k_femat(int nphase, int flag_m, const double *orient,
                                                  double c11, double c12, double c44,
                                                  double pm0, double pm1, double pm2, double pm3,
                                                  double *cmod, double *dk, double *sten,
                                                  float *dk32, float *sten32)
{
    // Every product and sum rounded on its own, in the order the reference writes them
    // (f90:2538-2627, 728-803): the tables come out bit for bit as its arithmetic gives
    // them.  (Fused multiply-adds change the last bits of cmod and dk; the initial energy
    // at 200^3, a difference of terms ~400 times larger than itself, then moves by ~1e-10.)
#pragma clang fp contract(off)
    __shared__ double cm[36];
    __shared__ double dks[576];
    const int p = blockIdx.x;
    if (threadIdx.x == 0) {
        double c[6][6] = {};
        const bool grain = flag_m == 1 && p < nphase - 1;
        if (grain) {
            // Rodrigues -> Q (f90:2538-2570)
            const double r0 = orient[3 * p], r1 = orient[3 * p + 1], r2 = orient[3 * p + 2];
            const double r[3] = {r0, r1, r2};
            double rdotr = r0 * r0 + r1 * r1 + r2 * r2;
            double t1 = 1.0 - rdotr, t2 = 1.0 + rdotr;
            double q[3][3];
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) q[i][j] = (i == j ? t1 : 0.0) + (2.0 * r[i]) * r[j];
            q[0][1] = q[0][1] - 2.0 * r2;
            q[0][2] = q[0][2] + 2.0 * r1;
            q[1][2] = q[1][2] - 2.0 * r0;
            q[1][0] = q[1][0] + 2.0 * r2;
            q[2][0] = q[2][0] - 2.0 * r1;
            q[2][1] = q[2][1] + 2.0 * r0;
            if (t2 > 1.0e-12) {
                t2 = 1.0 / t2;
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++) q[i][j] = q[i][j] * t2;
            } else {
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++) q[i][j] = i == j ? 1.0 : 0.0;
            }
            // Bond matrix newr (f90:2578-2621)
            double nr[6][6];
            for (int i = 0; i < 3; i++) {
                nr[i][0] = q[i][0] * q[i][0];
                nr[i][1] = q[i][1] * q[i][1];
                nr[i][2] = q[i][2] * q[i][2];
                nr[i][3] = q[i][1] * q[i][2];
                nr[i][4] = q[i][2] * q[i][0];
                nr[i][5] = q[i][0] * q[i][1];
            }
            // rows 4, 5, 6 pair (3,1), (2,3), (1,2) with engineering factors
            const int ra[3] = {2, 1, 0}, rb[3] = {0, 2, 1};
            for (int s = 0; s < 3; s++) {
                const int A = ra[s], B = rb[s];
                for (int k = 0; k < 3; k++) nr[3 + s][k] = 2.0 * q[A][k] * q[B][k];
            }
            nr[3][3] = q[0][1] * q[2][2] + q[0][2] * q[2][1];
            nr[3][4] = q[0][2] * q[2][0] + q[0][0] * q[2][2];
            nr[3][5] = q[0][0] * q[2][1] + q[0][1] * q[2][0];
            nr[4][3] = q[1][1] * q[2][2] + q[1][2] * q[2][1];
            nr[4][4] = q[1][0] * q[2][2] + q[1][2] * q[2][0];
            nr[4][5] = q[1][1] * q[2][0] + q[1][0] * q[2][1];
            nr[5][3] = q[0][1] * q[1][2] + q[0][2] * q[1][1];
            nr[5][4] = q[0][2] * q[1][0] + q[0][0] * q[1][2];
            nr[5][5] = q[0][0] * q[1][1] + q[0][1] * q[1][0];
            double cc[6][6] = {};
            cc[0][0] = cc[1][1] = cc[2][2] = c11;
            cc[0][1] = cc[0][2] = cc[1][2] = cc[1][0] = cc[2][0] = cc[2][1] = c12;
            cc[3][3] = cc[4][4] = cc[5][5] = c44;
            double t[6][6];
            for (int i = 0; i < 6; i++)
                for (int j = 0; j < 6; j++) {
                    double acc = 0.0;
                    for (int k = 0; k < 6; k++) acc += nr[k][i] * cc[k][j];
                    t[i][j] = acc;
                }
            for (int i = 0; i < 6; i++)
                for (int j = 0; j < 6; j++) {
                    double acc = 0.0;
                    for (int k = 0; k < 6; k++) acc += t[i][k] * nr[k][j];
                    c[i][j] = acc;
                }
        } else {
            // isotropic K*ck + G*cmu (f90:655-682, 700-705)
            double K, G;
            if (flag_m == 1) { K = pm0; G = pm1; }
            else if (p == 0) { K = pm0; G = pm1; }
            else { K = pm2; G = pm3; }
            const bool set = flag_m == 1 || p < 2;
            for (int i = 0; i < 6; i++)
                for (int j = 0; j < 6; j++) {
                    double ck = (i < 3 && j < 3) ? 1.0 : 0.0;
                    double cmu = (i < 3 && j < 3) ? (i == j ? 4.0 / 3.0 : -2.0 / 3.0) : (i == j ? 1.0 : 0.0);
                    c[i][j] = set ? K * ck + G * cmu : 0.0;
                }
        }
        for (int i = 0; i < 6; i++)
            for (int j = 0; j < 6; j++) {
                cm[i * 6 + j] = c[i][j];
                cmod[(int64_t)p * 36 + i * 6 + j] = c[i][j];
            }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < 576; e += kBlock) {
        const int j = e % 3, b = (e / 3) % 8, i = (e / 24) % 3, a = e / 72;
        double acc = 0.0;
        for (int k3 = 0; k3 < 3; k3++)
            for (int j3 = 0; j3 < 3; j3++)
                for (int i3 = 0; i3 < 3; i3++) {
                    const int nmid = (i3 == 1) + (j3 == 1) + (k3 == 1);
                    const double g = nmid == 0 ? 1.0 : (nmid == 1 ? 4.0 : (nmid == 2 ? 16.0 : 64.0));
                    const double x = i3 / 2.0, y = j3 / 2.0, z = k3 / 2.0;
                    double sumval = 0.0;
                    for (int r = 0; r < 6; r++) {
                        double tr = 0.0;
                        for (int q = 0; q < 6; q++) tr += cm[r * 6 + q] * es_entry(q, b, j, x, y, z);
                        sumval += es_entry(r, a, i, x, y, z) * tr;
                    }
                    acc += g * sumval / 216.0;
                }
        dks[e] = acc;
        dk[(int64_t)p * 576 + e] = acc;
        dk32[(int64_t)p * 576 + e] = (float)acc;
    }
    __syncthreads();
    __shared__ double sfull[243];
    for (int e = threadIdx.x; e < 243; e += kBlock) {
        const int n = e % 3, j = (e / 3) % 3, nb = e / 9;
        const int dx = nb % 3 - 1, dy = (nb / 3) % 3 - 1, dz = nb / 9 - 1;
        double acc = 0.0;
        bool first = true;
        for (int a = 0; a < 8; a++)
            for (int b = 0; b < 8; b++)
                if (PX(b) - PX(a) == dx && PY(b) - PY(a) == dy && PZ(b) - PZ(a) == dz) {
                    const double d = dks[((a * 3 + j) * 8 + b) * 3 + n];
                    acc = first ? d : acc + d;
                    first = false;
                }
        sfull[e] = acc;
    }
    __syncthreads();
    // half stencil: the 13 blocks of the neighbours nb = 14..26 (the block of -delta is
    // the transpose, K being symmetric) and the upper triangle of the centre block
    for (int e = threadIdx.x; e < kStenHalf; e += kBlock) {
        double v;
        if (e < 117) {
            v = sfull[(14 + e / 9) * 9 + e % 9];
        } else {
            const int c[6] = {0, 4, 8, 1, 2, 5};  // 00 11 22 01 02 12 of the centre block
            v = sfull[13 * 9 + c[e - 117]];
        }
        sten[(int64_t)p * kStenHalf + e] = v;
        sten32[(int64_t)p * kStenHalf + e] = (float)v;
    }
}

// ---------------------------------------------------------------------------
// Periodic-jump right-hand side (f90:815-1100), gathered per node so no atomics are
// needed: b(m, nn) = sum over the jump voxels v containing m (m is local a of v) of
// sum_{m3, m8} delta_v(m8, m3) dk(p_v, m8, m3, a, nn); and per voxel the constant
// C_v = 1/2 delta_v^T K_v delta_v, summed deterministically.
struct JumpArgs {
    int nx, ny, nz, nrhs;
    int64_t ns;
    const void *pix;  // phase ids, uint16 or (wide) uint32
    int wide;
    int64_t mstride;  // batch: nodes between the members' phase maps (0: one shared map)
    const double *dk;
    double e[6 * kMaxRhs];
};

__device__ __forceinline__ double jump_delta(const JumpArgs &A, int r, int fx, int fy, int fz,
                                             int a, int c)
{
    // sum in x, y, z order over the crossed upper faces of E_f = (e_fx, e_fy, e_fz) * n_f
    const double *e = A.e + 6 * r;  // exx eyy ezz exz eyz exy
    const double Ex[3] = {e[0] * A.nx, e[5] * A.nx, e[3] * A.nx};
    const double Ey[3] = {e[5] * A.ny, e[1] * A.ny, e[4] * A.ny};
    const double Ez[3] = {e[3] * A.nz, e[4] * A.nz, e[2] * A.nz};
    double s = 0.0;
    bool any = false;
    if (fx && PX(a)) { s = Ex[c]; any = true; }
    if (fy && PY(a)) { s = any ? s + Ey[c] : Ey[c]; any = true; }
    if (fz && PZ(a)) { s = any ? s + Ez[c] : Ez[c]; }
    return s;
}

// Only nodes on the grid's faces touch a voxel whose upper face is a periodic boundary
// (i, j or k at 0 or at the last index); the rest of b is zero.  Face node number e: the
// planes k = 0 and k = nz-1, then j = 0 and j = ny-1 with 0 < k < nz-1, then i = 0 and
// i = nx-1 with 0 < j < ny-1, 0 < k < nz-1 (each face node once, for extents >= 2).
__host__ __device__ inline int64_t face_count(int nx, int ny, int nz)
{
    return 2 * (int64_t)nx * ny + 2 * (int64_t)nx * (nz - 2) + 2 * (int64_t)(ny - 2) * (nz - 2);
}

__device__ __forceinline__ int64_t face_node(int64_t e, int nx, int ny, int nz)
{
    const int64_t nxy = (int64_t)nx * ny;
    if (e < 2 * nxy) return e % nxy + nxy * (e < nxy ? 0 : nz - 1);
    e -= 2 * nxy;
    const int64_t sj = (int64_t)nx * (nz - 2);
    if (e < 2 * sj) {
        const int j = e < sj ? 0 : ny - 1;
        const int64_t r = e % sj;
        return r % nx + nx * (j + (int64_t)ny * (1 + r / nx));
    }
    e -= 2 * sj;
    const int64_t si = (int64_t)(ny - 2) * (nz - 2);
    const int i = e < si ? 0 : nx - 1;
    const int64_t r = e % si;
    return i + nx * (1 + r % (ny - 2) + (int64_t)ny * (1 + r / (ny - 2)));
}

// b at the face nodes (the caller zeroes b first) and the constants of their voxels.
__global__ __launch_bounds__(kBlock) void k_rhs(JumpArgs A, double *b, double *partials)
{
    double cpart[kMaxRhs] = {};
    const int64_t nface = face_count(A.nx, A.ny, A.nz);
    for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < nface; e += (int64_t)gridDim.x * kBlock) {
        const int64_t m = face_node(e, A.nx, A.ny, A.nz);
        const int i = (int)(m % A.nx), j = (int)((m / A.nx) % A.ny), k = (int)(m / ((int64_t)A.nx * A.ny));
        for (int r = 0; r < A.nrhs; r++) {
            double acc[3] = {0.0, 0.0, 0.0};
            for (int a = 0; a < 8; a++) {
                const int vi = wrap(i - PX(a), A.nx), vj = wrap(j - PY(a), A.ny), vk = wrap(k - PZ(a), A.nz);
                const int fx = vi == A.nx - 1, fy = vj == A.ny - 1, fz = vk == A.nz - 1;
                if (!(fx | fy | fz)) continue;
                const int64_t v = vi + (int64_t)A.nx * (vj + (int64_t)A.ny * vk);
                const double *D = A.dk + (int64_t)phase_at(A.pix, A.wide, r * A.mstride + v) * 576;
                for (int nn = 0; nn < 3; nn++) {
                    double sv = 0.0;
                    for (int m3 = 0; m3 < 3; m3++)
                        for (int m8 = 0; m8 < 8; m8++)
                            sv = sv + jump_delta(A, r, fx, fy, fz, m8, m3) * D[((m8 * 3 + m3) * 8 + a) * 3 + nn];
                    acc[nn] += sv;
                }
            }
            for (int nn = 0; nn < 3; nn++) b[((int64_t)r * A.ns + m) * 3 + nn] = acc[nn];
            // constant of the voxel whose lower corner is m
            const int fx = i == A.nx - 1, fy = j == A.ny - 1, fz = k == A.nz - 1;
            if (fx | fy | fz) {
                const double *D = A.dk + (int64_t)phase_at(A.pix, A.wide, r * A.mstride + m) * 576;
                double c = 0.0;
                for (int nn = 0; nn < 3; nn++)
                    for (int mm = 0; mm < 8; mm++) {
                        const double dmn = jump_delta(A, r, fx, fy, fz, mm, nn);
                        for (int m3 = 0; m3 < 3; m3++)
                            for (int m8 = 0; m8 < 8; m8++)
                                c = c + 0.5 * jump_delta(A, r, fx, fy, fz, m8, m3) *
                                            D[((m8 * 3 + m3) * 8 + mm) * 3 + nn] * dmn;
                    }
                cpart[r] += c;
            }
        }
    }
    write_partials<kMaxRhs>(cpart, partials);
}

__global__ __launch_bounds__(kFinBlock) void k_rhs_finish(const double *partials, int nparts, int nrhs,
                                                       double *cconst)
{
    double tot[kMaxRhs];
    sum_partials<kMaxRhs>(partials, nparts, tot);
    if (threadIdx.x == 0)
        for (int r = 0; r < nrhs; r++) cconst[r] = tot[r];
}

// Affine initial guess u = E x (f90:423-438).
__global__ __launch_bounds__(kBlock) void k_affine(int nx, int ny, int nz, int nrhs, int64_t ns,
                                                   JumpArgs A, double *u)
{
#pragma clang fp contract(off)
    // separate multiply and add, as written in f90:432-434 (bitwise reproducible)
    const int64_t m = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (m >= ns) return;
    const double x = (double)(m % nx), y = (double)((m / nx) % ny), z = (double)(m / ((int64_t)nx * ny));
    for (int r = 0; r < nrhs; r++) {
        const double *e = A.e + 6 * r;
        const double exx = e[0], eyy = e[1], ezz = e[2], exz = e[3], eyz = e[4], exy = e[5];
        double *ur = u + ((int64_t)r * ns + m) * 3;
        ur[0] = x * exx + y * exy + z * exz;
        ur[1] = x * exy + y * eyy + z * eyz;
        ur[2] = x * exz + y * eyz + z * ezz;
    }
}

// ---------------------------------------------------------------------------
// Block-Jacobi preconditioner block of a node whose 8 voxels have phases ph (f90:1498-1520):
// M = sum over the voxels a of dk(ph[a], a, :, a, :), inverted by cofactors (invert3x3
// f90:2046-2147); o = ainv(1,1), (2,2), (3,3), (1,2), (1,3), (2,3).  One function for the
// per-node array and the per-phase table of homogeneous nodes, without contraction, so
// that both give the same bits.
__device__ __forceinline__ void precond_block(const double *dk, const int (&ph)[8], double (&o)[6])
{
#pragma clang fp contract(off)
    double M[3][3];
    for (int a = 0; a < 8; a++) {
        const double *D = dk + (int64_t)ph[a] * 576;
        for (int ii = 0; ii < 3; ii++)
            for (int jj = 0; jj < 3; jj++) {
                const double d = D[((a * 3 + ii) * 8 + a) * 3 + jj];
                M[ii][jj] = a == 0 ? d : M[ii][jj] + d;
            }
    }
    const double c11 = M[1][1] * M[2][2] - M[1][2] * M[2][1];
    const double c21 = -M[1][0] * M[2][2] + M[1][2] * M[2][0];
    const double c31 = M[1][0] * M[2][1] - M[1][1] * M[2][0];
    const double det = M[0][0] * c11 + M[0][1] * c21 + M[0][2] * c31;
    for (int q = 0; q < 6; q++) o[q] = 0.0;
    if (!(fabs(det) < 1e-14)) {
        const double inv = 1.0 / det;
        const double c12 = -M[0][1] * M[2][2] + M[0][2] * M[2][1];
        const double c22 = M[0][0] * M[2][2] - M[0][2] * M[2][0];
        const double c13 = M[0][1] * M[1][2] - M[0][2] * M[1][1];
        const double c23 = -M[0][0] * M[1][2] + M[0][2] * M[1][0];
        const double c33 = M[0][0] * M[1][1] - M[0][1] * M[1][0];
        // ainv(1,1), (2,2), (3,3), (1,2), (1,3), (2,3)
        o[0] = c11 * inv; o[1] = c22 * inv; o[2] = c33 * inv;
        o[3] = c12 * inv; o[4] = c13 * inv; o[5] = c23 * inv;
    }
}

// Per node: M^-1 (and its FP32 copy), and hphase[m] = the phase of its 8 voxels when they
// share one, else the all-ones id of the handle's width; the streaming update takes homogeneous nodes' M^-1 from the
// per-phase table (k_precond_table) and reads this array only for the others.  identity:
// the plain-CG loop's M^-1 = I (z = 1 g0 + 0 g1 + 0 g2 = g exactly for finite g).
__global__ __launch_bounds__(kBlock) void k_precond(int nx, int ny, int nz, int64_t ns,
                                                    const void *pix, int wide, const double *dk,
                                                    double *minv, float *minv32, void *hphase, int identity)
{
    const int64_t m = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (m >= ns) return;
    const int i = (int)(m % nx), j = (int)((m / nx) % ny), k = (int)(m / ((int64_t)nx * ny));
    int ph[8];
    bool homog = true;
    for (int a = 0; a < 8; a++) {
        const int64_t v = wrap(i - PX(a), nx) + (int64_t)nx * (wrap(j - PY(a), ny) + (int64_t)ny * wrap(k - PZ(a), nz));
        ph[a] = phase_at(pix, wide, v);
        homog &= ph[a] == ph[0];
    }
    if (wide) static_cast<uint32_t *>(hphase)[m] = homog ? (uint32_t)ph[0] : 0xFFFFFFFFu;
    else static_cast<uint16_t *>(hphase)[m] = homog ? (uint16_t)ph[0] : (uint16_t)0xFFFF;
    double o[6];
    if (identity) {
        for (int q = 0; q < 6; q++) o[q] = q < 3 ? 1.0 : 0.0;
    } else {
        precond_block(dk, ph, o);
    }
    for (int q = 0; q < 6; q++) {
        minv[6 * m + q] = o[q];
        if (minv32) minv32[6 * m + q] = (float)o[q];
    }
}

// M^-1 of a homogeneous node of each phase, [nphase][6] (the same function as k_precond).
__global__ __launch_bounds__(64) void k_precond_table(int nphase, const double *dk, double *tab, float *tab32,
                                                      int identity)
{
    const int p = blockIdx.x * 64 + threadIdx.x;
    if (p >= nphase) return;
    int ph[8];
    for (int a = 0; a < 8; a++) ph[a] = p;
    double o[6];
    if (identity) {
        for (int q = 0; q < 6; q++) o[q] = q < 3 ? 1.0 : 0.0;
    } else {
        precond_block(dk, ph, o);
    }
    for (int q = 0; q < 6; q++) {
        tab[6 * p + q] = o[q];
        if (tab32) tab32[6 * p + q] = (float)o[q];
    }
}

JumpArgs jump_args(const Solver &s)
{
    JumpArgs A{};
    A.nx = s.nx; A.ny = s.ny; A.nz = s.nz; A.nrhs = s.nrhs; A.ns = s.ns;
    A.pix = s.pix; A.wide = s.wide; A.mstride = s.mstride(); A.dk = s.dk;
    for (int q = 0; q < 6 * kMaxRhs; q++) A.e[q] = s.strain[q];
    return A;
}

}  // namespace

hipError_t launch_femat(Solver &s, double c11, double c12, double c44, const double *pm)
{
    return launch_femat_range(s, 0, s.nphase, s.orient, c11, c12, c44, pm);
}

hipError_t launch_femat_range(Solver &s, int base, int nphase, const double *orient, double c11, double c12,
                              double c44, const double *pm)
{
    const int64_t b = base;
    hipLaunchKernelGGL(k_femat, dim3(nphase), dim3(kBlock), 0, s.stream, nphase, s.flag_m, orient, c11, c12, c44,
                       pm[0], pm[1], pm[2], pm[3], s.cmod + 36 * b, s.dk + 576 * b, s.sten + kStenHalf * b,
                       s.dk32 + 576 * b, s.sten32 + kStenHalf * b);
    return hipGetLastError();
}

hipError_t launch_rhs(Solver &s)
{
    const int64_t nface = face_count(s.nx, s.ny, s.nz);
    const int g = (int)std::min<int64_t>(vector_grid(s), (nface + kBlock - 1) / kBlock);
    hipError_t e = hipMemsetAsync(s.b, 0, sizeof(double) * 3 * (size_t)s.nrhs * (size_t)s.ns, s.stream);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_rhs, dim3(g), dim3(kBlock), 0, s.stream, jump_args(s), s.b, s.partials);
    hipLaunchKernelGGL(k_rhs_finish, dim3(1), dim3(kFinBlock), 0, s.stream, (const double *)s.partials, g, s.nrhs,
                       s.cconst);
    return hipGetLastError();
}

hipError_t launch_affine(Solver &s)
{
    hipLaunchKernelGGL(k_affine, dim3((unsigned)vector_blocks(s)), dim3(kBlock), 0, s.stream, s.nx, s.ny,
                       s.nz, s.nrhs, s.ns, jump_args(s), s.u);
    return hipGetLastError();
}

hipError_t launch_precond(Solver &s, bool want32)
{
    const int identity = s.pcg_variant == 2 ? 1 : 0;
    const int64_t idb = s.wide ? 4 : 2;  // phase-id bytes
    for (int64_t mb = 0; mb < s.member_count(); mb++) {  // each member's microstructure
        const int64_t n0 = mb * s.ns;
        hipLaunchKernelGGL(k_precond, dim3((unsigned)vector_blocks(s)), dim3(kBlock), 0, s.stream, s.nx, s.ny,
                           s.nz, s.ns, (const void *)((const char *)s.pix + idb * n0), s.wide ? 1 : 0,
                           (const double *)s.dk, s.minv + 6 * n0, want32 ? s.minv32 + 6 * n0 : nullptr,
                           (void *)((char *)s.hphase + idb * n0), identity);
    }
    hipLaunchKernelGGL(k_precond_table, dim3((unsigned)((s.nphase + 63) / 64)), dim3(64), 0, s.stream, s.nphase,
                       (const double *)s.dk, s.minv_hom, want32 ? s.minv_hom32 : nullptr, identity);
    s.precond_identity = s.pcg_variant == 2;
    return hipGetLastError();
}

}  // namespace e3x
