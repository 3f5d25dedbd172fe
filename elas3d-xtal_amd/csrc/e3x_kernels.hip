// gfx950 kernels of the elas3dxtal_pcg solve.
//
// Layout in HBM (structure of arrays, Fortran column-major as in the reference):
//   fields  [rhs][component][node], node m = i + nx*(j + ny*k)
//   pix     uint16 0-based phase of the voxel whose lower corner is node m
//   tables  cmod[p][6][6], dk[p][a][i][b][j] (576), sten[p][27][3][3] (243)
//   minv    6 planes (00, 11, 22, 01, 02, 12) of the symmetric block inverse
// Reductions are deterministic: fixed-shape wave/block trees, per-block partials,
// and a fixed-order final sum by the last block to arrive.
#include "e3x_internal.hpp"

#include <cmath>

namespace e3x {
namespace {

// Local node a of the brick sits at (PX(a), PY(a), PZ(a)):
// 1 (0,0,0) 2 (1,0,0) 3 (1,1,0) 4 (0,1,0) 5 (0,0,1) 6 (1,0,1) 7 (1,1,1) 8 (0,1,1)
// (src/elas3dxtal_pcg.f90:639, 740-765).
__host__ __device__ constexpr int PX(int a) { return ((a + 1) >> 1) & 1; }
__host__ __device__ constexpr int PY(int a) { return (a >> 1) & 1; }
__host__ __device__ constexpr int PZ(int a) { return a >> 2; }

__device__ __forceinline__ int wrap(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// Consecutive tile ids on one XCD: hardware hands block b to XCD b % 8.
__device__ __forceinline__ int xcd_remap(int bid, int nb)
{
    int xcd = bid % kXcds, k = bid / kXcds;
    int q = nb / kXcds, rem = nb % kXcds;
    return xcd * q + (xcd < rem ? xcd : rem) + k;
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// Deterministic sum over the block of NV values; the result is valid in thread 0.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double (*sh)[kBlock / 64])
{
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
    for (int q = 0; q < NV; q++) {
        double t = wave_sum(v[q]);
        if (l == 0) sh[q][w] = t;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; q++) {
            double t = sh[q][0];
            for (int ww = 1; ww < kBlock / 64; ww++) t += sh[q][ww];
            v[q] = t;
        }
    }
}

__device__ __forceinline__ double load_coherent(const double *p)
{
    return __hip_atomic_load(const_cast<double *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Write this block's NV partials (layout [q][block]) and report whether this block
// is the last one to arrive; in that block `tot` receives the fixed-order totals.
template <int NV>
__device__ bool reduce_partials(double (&v)[NV], double *partials, unsigned *counter,
                                double (&tot)[NV])
{
    __shared__ double sh[NV > 0 ? NV : 1][kBlock / 64];
    __shared__ int is_last;
    const int nb = gridDim.x;
    block_sum<NV>(v, sh);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; q++) partials[(int64_t)q * nb + blockIdx.x] = v[q];
        __threadfence();
        unsigned prev = atomicAdd(counter, 1u);
        is_last = (prev == (unsigned)(nb - 1));
    }
    __syncthreads();
    if (!is_last) return false;
    __threadfence();
    double acc[NV];
#pragma unroll
    for (int q = 0; q < NV; q++) {
        double t = 0.0;
        for (int i = threadIdx.x; i < nb; i += kBlock) t += load_coherent(&partials[(int64_t)q * nb + i]);
        acc[q] = t;
    }
    __syncthreads();
    block_sum<NV>(acc, sh);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; q++) tot[q] = acc[q];
        *counter = 0u;
    }
    return true;
}

// ---------------------------------------------------------------------------
// femat: one workgroup per phase (rotatestiffnessby f90:2470-2637, cmod f90:655-706,
// dk by 3x3x3 Simpson f90:713-805) plus the homogeneous 27-point stencil blocks
// sten[p][nb][j][n] = sum over the (voxel, local a, local b) pairs that couple a node
// to its neighbour nb of dk[p][a][j][b][n].
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
        sten[(int64_t)p * 243 + e] = acc;
        sten32[(int64_t)p * 243 + e] = (float)acc;
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
    const uint16_t *pix;
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

__global__ __launch_bounds__(kBlock) void k_rhs(JumpArgs A, double *b, double *partials,
                                                unsigned *counter, double *cconst)
{
    const int64_t m = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    double cpart[kMaxRhs] = {};
    if (m < A.ns) {
        const int i = (int)(m % A.nx), j = (int)((m / A.nx) % A.ny), k = (int)(m / ((int64_t)A.nx * A.ny));
        for (int r = 0; r < A.nrhs; r++) {
            double acc[3] = {0.0, 0.0, 0.0};
            for (int a = 0; a < 8; a++) {
                const int vi = wrap(i - PX(a), A.nx), vj = wrap(j - PY(a), A.ny), vk = wrap(k - PZ(a), A.nz);
                const int fx = vi == A.nx - 1, fy = vj == A.ny - 1, fz = vk == A.nz - 1;
                if (!(fx | fy | fz)) continue;
                const int64_t v = vi + (int64_t)A.nx * (vj + (int64_t)A.ny * vk);
                const double *D = A.dk + (int64_t)A.pix[v] * 576;
                for (int nn = 0; nn < 3; nn++) {
                    double sv = 0.0;
                    for (int m3 = 0; m3 < 3; m3++)
                        for (int m8 = 0; m8 < 8; m8++)
                            sv = sv + jump_delta(A, r, fx, fy, fz, m8, m3) * D[((m8 * 3 + m3) * 8 + a) * 3 + nn];
                    acc[nn] += sv;
                }
            }
            for (int nn = 0; nn < 3; nn++) b[((int64_t)r * 3 + nn) * A.ns + m] = acc[nn];
            // constant of the voxel whose lower corner is m
            const int fx = i == A.nx - 1, fy = j == A.ny - 1, fz = k == A.nz - 1;
            if (fx | fy | fz) {
                const double *D = A.dk + (int64_t)A.pix[m] * 576;
                double c = 0.0;
                for (int nn = 0; nn < 3; nn++)
                    for (int mm = 0; mm < 8; mm++) {
                        const double dmn = jump_delta(A, r, fx, fy, fz, mm, nn);
                        for (int m3 = 0; m3 < 3; m3++)
                            for (int m8 = 0; m8 < 8; m8++)
                                c = c + 0.5 * jump_delta(A, r, fx, fy, fz, m8, m3) *
                                            D[((m8 * 3 + m3) * 8 + mm) * 3 + nn] * dmn;
                    }
                cpart[r] = c;
            }
        }
    }
    double tot[kMaxRhs];
    if (reduce_partials<kMaxRhs>(cpart, partials, counter, tot) && threadIdx.x == 0)
        for (int r = 0; r < A.nrhs; r++) cconst[r] = tot[r];
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
        double *ur = u + (int64_t)r * 3 * ns;
        ur[m] = x * exx + y * exy + z * exz;
        ur[ns + m] = x * exy + y * eyy + z * eyz;
        ur[2 * ns + m] = x * exz + y * eyz + z * ezz;
    }
}

// ---------------------------------------------------------------------------
// Matrix-free stencil y = K x (energy_OpenMP f90:1176-1321, dembx SpMV f90:1585-1884).
//
// A block stages its 10x10x6 node halo of x and 9x9x5 voxel phases in LDS.  A wave
// whose 64 nodes all sit inside one phase applies that phase's 27-point stencil with
// wave-uniform (scalar-cache) coefficients: 243 FMA per node per right-hand side.
// Any other wave uses the general element form: for each of the node's 8 voxels, the
// 3x24 rows of dk for the node's local index: 576 FMA.
template <typename T>
struct StencilArgs {
    int nx, ny, nz, ntx, nty;
    int64_t ns;
    const uint16_t *pix;
    const T *dk, *sten;
    const T *x;            // [NR][3][ns]
    T *y;                  // [NR][3][ns] (kApply, kEnergy, kPlain)
    const double *b;       // kEnergy, kResid
    float *y32;            // kResid output
    double *partials;
    unsigned *counter;
    PcgState *st;          // kApply
    double *results;       // kEnergy: utot (C added) then gg per rhs
    const double *cconst;  // kEnergy
    int it;
};

template <typename T, int NR, int MODE>
__global__ __launch_bounds__(kBlock) void k_stencil(StencilArgs<T> A)
{
    __shared__ T hs[NR * 3 * HN];
    __shared__ int ps[VN];

    bool active_r[NR];
    bool any = true;
    if constexpr (MODE == kApply) {
        any = false;
#pragma unroll
        for (int r = 0; r < NR; r++) {
            active_r[r] = A.st[r].stop_iter == 0;
            any |= active_r[r];
        }
        if (!any) return;
    }

    const int bid = xcd_remap(blockIdx.x, gridDim.x);
    const int tx = bid % A.ntx, ty = (bid / A.ntx) % A.nty, tz = bid / (A.ntx * A.nty);
    const int ox = tx * TX, oy = ty * TY, oz = tz * TZ;
    const int64_t ns = A.ns;

    for (int q = threadIdx.x; q < HN; q += kBlock) {
        const int hx = q % HX, hy = (q / HX) % HY, hz = q / (HX * HY);
        const int gx = wrap(ox + hx - 1, A.nx), gy = wrap(oy + hy - 1, A.ny), gz = wrap(oz + hz - 1, A.nz);
        const int64_t gm = gx + (int64_t)A.nx * (gy + (int64_t)A.ny * gz);
#pragma unroll
        for (int rc = 0; rc < NR * 3; rc++) hs[rc * HN + q] = A.x[(int64_t)rc * ns + gm];
    }
    for (int q = threadIdx.x; q < VN; q += kBlock) {
        const int vx = q % VX, vy = (q / VX) % VY, vz = q / (VX * VY);
        const int gx = wrap(ox + vx - 1, A.nx), gy = wrap(oy + vy - 1, A.ny), gz = wrap(oz + vz - 1, A.nz);
        ps[q] = A.pix[gx + (int64_t)A.nx * (gy + (int64_t)A.ny * gz)];
    }
    __syncthreads();

    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int lx = (w & 1) * 4 + (l & 3), ly = (w >> 1) * 4 + ((l >> 2) & 3), lz = l >> 4;
    const int gx = ox + lx, gy = oy + ly, gz = oz + lz;
    const bool active = gx < A.nx && gy < A.ny && gz < A.nz;
    const int64_t m = gx + (int64_t)A.nx * (gy + (int64_t)A.ny * gz);

    int pv[8];
#pragma unroll
    for (int a = 0; a < 8; a++) pv[a] = ps[(lz + 1 - PZ(a)) * (VX * VY) + (ly + 1 - PY(a)) * VX + (lx + 1 - PX(a))];
    bool homog = true;
#pragma unroll
    for (int a = 1; a < 8; a++) homog &= pv[a] == pv[0];

    const unsigned long long am = __ballot(active);
    const int first = am ? __ffsll((long long)am) - 1 : 0;
    const int pref = __shfl(pv[0], first, 64);
    const bool ok = !active || (homog && pv[0] == pref);
    const bool uniform = __ballot(ok) == ~0ull;

    const int cidx = (lz + 1) * (HX * HY) + (ly + 1) * HX + (lx + 1);
    T acc[NR][3];
#pragma unroll
    for (int r = 0; r < NR; r++) acc[r][0] = acc[r][1] = acc[r][2] = T(0);

    if (am != 0ull) {
        if (uniform) {
            const int P = __builtin_amdgcn_readfirstlane(pref);
            const T *S = A.sten + (int64_t)P * 243;
#pragma unroll
            for (int nb = 0; nb < 27; nb++) {
                const int dx = nb % 3 - 1, dy = (nb / 3) % 3 - 1, dz = nb / 9 - 1;
                const int hidx = cidx + dz * (HX * HY) + dy * HX + dx;
#pragma unroll
                for (int r = 0; r < NR; r++) {
                    const T h0 = hs[(r * 3 + 0) * HN + hidx], h1 = hs[(r * 3 + 1) * HN + hidx],
                            h2 = hs[(r * 3 + 2) * HN + hidx];
#pragma unroll
                    for (int j = 0; j < 3; j++) {
                        acc[r][j] += S[nb * 9 + j * 3 + 0] * h0;
                        acc[r][j] += S[nb * 9 + j * 3 + 1] * h1;
                        acc[r][j] += S[nb * 9 + j * 3 + 2] * h2;
                    }
                }
            }
        } else {
#pragma unroll
            for (int a = 0; a < 8; a++) {
                const T *D = A.dk + (int64_t)pv[a] * 576 + a * 72;
#pragma unroll
                for (int b = 0; b < 8; b++) {
                    const int hidx = cidx + (PZ(b) - PZ(a)) * (HX * HY) + (PY(b) - PY(a)) * HX + (PX(b) - PX(a));
#pragma unroll
                    for (int r = 0; r < NR; r++) {
                        const T h0 = hs[(r * 3 + 0) * HN + hidx], h1 = hs[(r * 3 + 1) * HN + hidx],
                                h2 = hs[(r * 3 + 2) * HN + hidx];
#pragma unroll
                        for (int j = 0; j < 3; j++) {
                            acc[r][j] += D[j * 24 + b * 3 + 0] * h0;
                            acc[r][j] += D[j * 24 + b * 3 + 1] * h1;
                            acc[r][j] += D[j * 24 + b * 3 + 2] * h2;
                        }
                    }
                }
            }
        }
    }

    if constexpr (MODE == kApply) {
        double part[NR];
#pragma unroll
        for (int r = 0; r < NR; r++) {
            part[r] = 0.0;
            if (active) {
#pragma unroll
                for (int j = 0; j < 3; j++) {
                    A.y[((int64_t)r * 3 + j) * ns + m] = acc[r][j];
                    part[r] += (double)hs[(r * 3 + j) * HN + cidx] * (double)acc[r][j];
                }
            }
        }
        double tot[NR];
        if (reduce_partials<NR>(part, A.partials, A.counter, tot) && threadIdx.x == 0) {
            // f90:1889-1907
            for (int r = 0; r < NR; r++) {
                PcgState &S = A.st[r];
                if (!active_r[r]) continue;
                S.hah = tot[r];
                if (fabs(tot[r]) < 1e-14) {
                    S.stop_iter = A.it;
                    S.exit_flag = 2;  // E3X_PCG_HAH_BREAK
                    S.run_dir = 0;
                } else {
                    S.lambda = S.gg / tot[r];
                }
            }
        }
    } else if constexpr (MODE == kEnergy) {
        double part[2 * NR];
#pragma unroll
        for (int r = 0; r < NR; r++) {
            double e = 0.0, gsq = 0.0;
            if (active) {
#pragma unroll
                for (int j = 0; j < 3; j++) {
                    const int64_t q = ((int64_t)r * 3 + j) * ns + m;
                    const double uq = (double)hs[(r * 3 + j) * HN + cidx];
                    const double bq = A.b[q];
                    const double kx = (double)acc[r][j];
                    e = e + 0.5 * uq * kx + bq * uq;  // f90:1336
                    const double gq = kx + bq;         // f90:1337
                    A.y[q] = (T)gq;
                    gsq += gq * gq;                     // f90:452
                }
            }
            part[r] = e;
            part[NR + r] = gsq;
        }
        double tot[2 * NR];
        if (reduce_partials<2 * NR>(part, A.partials, A.counter, tot) && threadIdx.x == 0) {
            for (int r = 0; r < NR; r++) {
                A.results[r] = A.cconst[r] + tot[r];
                A.results[NR + r] = tot[NR + r];
            }
        }
    } else if constexpr (MODE == kResid) {
        if (active) {
#pragma unroll
            for (int r = 0; r < NR; r++) {
                if (A.st[r].stop_iter != 0) continue;
#pragma unroll
                for (int j = 0; j < 3; j++) {
                    const int64_t q = ((int64_t)r * 3 + j) * ns + m;
                    A.y32[q] = (float)((double)acc[r][j] + A.b[q]);
                }
            }
        }
    } else {
        if (active) {
#pragma unroll
            for (int r = 0; r < NR; r++)
#pragma unroll
                for (int j = 0; j < 3; j++) A.y[((int64_t)r * 3 + j) * ns + m] = acc[r][j];
        }
    }
}

// ---------------------------------------------------------------------------
// Block-Jacobi preconditioner (f90:1498-1520): M = sum over the 8 voxels of the node
// of dk(p, a, :, a, :), inverted by cofactors (invert3x3 f90:2046-2147).
__global__ __launch_bounds__(kBlock) void k_precond(int nx, int ny, int nz, int64_t ns,
                                                    const uint16_t *pix, const double *dk,
                                                    double *minv, float *minv32)
{
    const int64_t m = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (m >= ns) return;
    const int i = (int)(m % nx), j = (int)((m / nx) % ny), k = (int)(m / ((int64_t)nx * ny));
    double M[3][3];
    for (int a = 0; a < 8; a++) {
        const int64_t v = wrap(i - PX(a), nx) + (int64_t)nx * (wrap(j - PY(a), ny) + (int64_t)ny * wrap(k - PZ(a), nz));
        const double *D = dk + (int64_t)pix[v] * 576;
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
    double o[6] = {0, 0, 0, 0, 0, 0};
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
    for (int q = 0; q < 6; q++) {
        minv[(int64_t)q * ns + m] = o[q];
        if (minv32) minv32[(int64_t)q * ns + m] = (float)o[q];
    }
}

// ---------------------------------------------------------------------------
// PCG vector kernels.  One node per thread; T is the storage type of g, h, z, Ah and
// the preconditioner (double, or float in the mixed form); u is always double and
// every dot product is accumulated in double.
template <typename T>
struct PcgArgs {
    int64_t ns;
    int it, ldemb, kkk;
    double *u;
    T *g, *h, *z, *ah;
    const T *minv;
    double *partials;
    unsigned *counter;
    PcgState *st;
    double *hist;
};

template <typename T>
__device__ __forceinline__ void precond_apply(const T *minv, int64_t ns, int64_t m, double g0,
                                              double g1, double g2, double &z0, double &z1,
                                              double &z2)
{
    const double a00 = minv[m], a11 = minv[ns + m], a22 = minv[2 * ns + m];
    const double a01 = minv[3 * ns + m], a02 = minv[4 * ns + m], a12 = minv[5 * ns + m];
    z0 = a00 * g0 + a01 * g1 + a02 * g2;
    z1 = a01 * g0 + a11 * g1 + a12 * g2;
    z2 = a02 * g0 + a12 * g1 + a22 * g2;
}

// z = M^-1 g, gg = z.g, (kkk == 1) h = z; then the loop-entry scalars (f90:1527-1567).
template <typename T, int NR>
__global__ __launch_bounds__(kBlock) void k_pcg_init(PcgArgs<T> A)
{
    const int64_t m = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t ns = A.ns;
    double part[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) {
        part[r] = 0.0;
        if (m < ns) {
            T *g = A.g + (int64_t)r * 3 * ns, *z = A.z + (int64_t)r * 3 * ns, *h = A.h + (int64_t)r * 3 * ns;
            const double g0 = g[m], g1 = g[ns + m], g2 = g[2 * ns + m];
            double z0, z1, z2;
            precond_apply(A.minv, ns, m, g0, g1, g2, z0, z1, z2);
            z[m] = (T)z0; z[ns + m] = (T)z1; z[2 * ns + m] = (T)z2;
            if (A.kkk == 1) { h[m] = (T)z0; h[ns + m] = (T)z1; h[2 * ns + m] = (T)z2; }
            part[r] = (double)(T)z0 * g0 + (double)(T)z1 * g1 + (double)(T)z2 * g2;
        }
    }
    double tot[NR];
    if (reduce_partials<NR>(part, A.partials, A.counter, tot) && threadIdx.x == 0) {
        for (int r = 0; r < NR; r++) {
            PcgState &S = A.st[r];
            S.gg = tot[r];
            S.rel_res_prev = sqrt(S.gg / S.gg0);
            S.rel_res_min = S.rel_res_prev;
            S.rel_res = S.rel_res_prev;
            S.reset_count = 0;
            S.reset = 0;
            S.run_dir = 0;
            if (A.ldemb <= 0) {
                S.stop_iter = -1;
                S.exit_flag = 3;
            } else {
                S.stop_iter = 0;
                S.exit_flag = 0;
            }
        }
    }
}

// Scalar bookkeeping after z.g (f90:1938-2007).
__device__ void finish_update(PcgState &S, double gg_new, int it, double *hist, int r)
{
    const double gglast = S.gg;
    S.gg = gg_new;
    const double rel = sqrt(S.gg / S.gg0);
    S.rel_res = rel;
    hist[(int64_t)r * S.ldemb + it - 1] = rel;
    if (rel > 10.0 * S.rel_res_prev && it > 1) {
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
    if (rel < S.tol) {
        S.stop_iter = it;
        S.exit_flag = 1;
        S.run_dir = 1;
    } else if (it >= S.ldemb) {
        S.stop_iter = it;
        S.exit_flag = 3;
        S.run_dir = 1;
    }
}

// u -= lambda h; g -= lambda Ah (unless the true gradient replaces g this iteration);
// z = M^-1 g; gg = z.g   (f90:1914-1946)
template <typename T, int NR, bool UPDATE_G, bool DO_U, bool DO_Z>
__global__ __launch_bounds__(kBlock) void k_pcg_update(PcgArgs<T> A)
{
    bool act[NR];
    bool any = false;
#pragma unroll
    for (int r = 0; r < NR; r++) {
        act[r] = A.st[r].stop_iter == 0;
        any |= act[r];
    }
    if (!any) return;
    const int64_t m = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t ns = A.ns;
    double part[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) {
        part[r] = 0.0;
        if (m >= ns || !act[r]) continue;
        const int64_t o = (int64_t)r * 3 * ns + m;
        const double lam = A.st[r].lambda;
        if constexpr (DO_U) {
#pragma unroll
            for (int c = 0; c < 3; c++) A.u[o + c * ns] = A.u[o + c * ns] - lam * (double)A.h[o + c * ns];
        }
        double g[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            g[c] = (double)A.g[o + c * ns];
            if constexpr (UPDATE_G) {
                g[c] = g[c] - lam * (double)A.ah[o + c * ns];
                A.g[o + c * ns] = (T)g[c];
                g[c] = (double)(T)g[c];
            }
        }
        if constexpr (DO_Z) {
            double z0, z1, z2;
            precond_apply(A.minv, ns, m, g[0], g[1], g[2], z0, z1, z2);
            const T t0 = (T)z0, t1 = (T)z1, t2 = (T)z2;
            A.z[o] = t0; A.z[o + ns] = t1; A.z[o + 2 * ns] = t2;
            part[r] = (double)t0 * g[0] + (double)t1 * g[1] + (double)t2 * g[2];
        }
    }
    if constexpr (DO_Z) {
        double tot[NR];
        if (reduce_partials<NR>(part, A.partials, A.counter, tot) && threadIdx.x == 0) {
            for (int r = 0; r < NR; r++)
                if (act[r]) finish_update(A.st[r], tot[r], A.it, A.hist, r);
        }
    }
}

// h = z + gamma h, or h = z on a reset (f90:1961-1996).
template <typename T, int NR>
__global__ __launch_bounds__(kBlock) void k_pcg_direction(PcgArgs<T> A)
{
    const int64_t m = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (m >= A.ns) return;
    const int64_t ns = A.ns;
#pragma unroll
    for (int r = 0; r < NR; r++) {
        const PcgState &S = A.st[r];
        const bool run = S.stop_iter == 0 || (S.stop_iter == A.it && S.run_dir);
        if (!run) continue;
        const int64_t o = (int64_t)r * 3 * ns + m;
        if (S.reset) {
#pragma unroll
            for (int c = 0; c < 3; c++) A.h[o + c * ns] = A.z[o + c * ns];
        } else {
            const double gam = S.gamma;
#pragma unroll
            for (int c = 0; c < 3; c++)
                A.h[o + c * ns] = (T)((double)A.z[o + c * ns] + gam * (double)A.h[o + c * ns]);
        }
    }
}

// ---------------------------------------------------------------------------
// Voxel-centre strain and stress (stress_OpenMP f90:2149-2408,
// stress_fullfield_OpenMP f90:2639-2833).  es at the centre is +-1/4.
struct PostArgs {
    int nx, ny, nz, nrhs;
    int64_t ns;
    const uint16_t *pix;
    const double *cmod, *u;
    double e[6 * kMaxRhs];
};

__device__ void voxel_strain_stress(const PostArgs &A, int r, int64_t m, double eps[6], double sig[6])
{
    const int i = (int)(m % A.nx), j = (int)((m / A.nx) % A.ny), k = (int)(m / ((int64_t)A.nx * A.ny));
    const double *u = A.u + (int64_t)r * 3 * A.ns;
    const double *e = A.e + 6 * r;  // exx eyy ezz exz eyz exy
    double uu[8][3];
    for (int a = 0; a < 8; a++) {
        const int64_t q = wrap(i + PX(a), A.nx) + (int64_t)A.nx * (wrap(j + PY(a), A.ny) + (int64_t)A.ny * wrap(k + PZ(a), A.nz));
        for (int c = 0; c < 3; c++) uu[a][c] = u[c * A.ns + q];
    }
    if (i == A.nx - 1)
        for (int a = 0; a < 8; a++)
            if (PX(a)) { uu[a][0] += e[0] * A.nx; uu[a][1] += e[5] * A.nx; uu[a][2] += e[3] * A.nx; }
    if (j == A.ny - 1)
        for (int a = 0; a < 8; a++)
            if (PY(a)) { uu[a][0] += e[5] * A.ny; uu[a][1] += e[1] * A.ny; uu[a][2] += e[4] * A.ny; }
    if (k == A.nz - 1)
        for (int a = 0; a < 8; a++)
            if (PZ(a)) { uu[a][0] += e[3] * A.nz; uu[a][1] += e[4] * A.nz; uu[a][2] += e[2] * A.nz; }
    for (int q = 0; q < 6; q++) eps[q] = 0.0;
    for (int n3 = 0; n3 < 3; n3++)
        for (int a = 0; a < 8; a++) {
            const double dx = PX(a) ? 0.25 : -0.25, dy = PY(a) ? 0.25 : -0.25, dz = PZ(a) ? 0.25 : -0.25;
            const double v = uu[a][n3];
            if (n3 == 0) { eps[0] += dx * v; eps[3] += dz * v; eps[5] += dy * v; }
            else if (n3 == 1) { eps[1] += dy * v; eps[4] += dz * v; eps[5] += dx * v; }
            else { eps[2] += dz * v; eps[3] += dx * v; eps[4] += dy * v; }
        }
    const double *c = A.cmod + (int64_t)A.pix[m] * 36;
    for (int q = 0; q < 6; q++) {
        double s = 0.0;
        for (int t = 0; t < 6; t++) s += c[q * 6 + t] * eps[t];
        sig[q] = s;
    }
}

template <int NR>
__global__ __launch_bounds__(kBlock) void k_stress(PostArgs A, double *partials, unsigned *counter,
                                                   double *results)
{
    const int64_t m = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    double part[12 * NR];
#pragma unroll
    for (int q = 0; q < 12 * NR; q++) part[q] = 0.0;
    if (m < A.ns) {
#pragma unroll
        for (int r = 0; r < NR; r++) {
            double eps[6], sig[6];
            voxel_strain_stress(A, r, m, eps, sig);
#pragma unroll
            for (int q = 0; q < 6; q++) {
                part[12 * r + q] = sig[q];
                part[12 * r + 6 + q] = eps[q];
            }
        }
    }
    double tot[12 * NR];
    if (reduce_partials<12 * NR>(part, partials, counter, tot) && threadIdx.x == 0)
        for (int q = 0; q < 12 * NR; q++) results[q] = tot[q] / (double)A.ns;
}

__global__ __launch_bounds__(kBlock) void k_fullfield(PostArgs A, int r, double *sf, double *vm)
{
    const int64_t m = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (m >= A.ns) return;
    double eps[6], s[6];
    voxel_strain_stress(A, r, m, eps, s);
    if (sf)
        for (int q = 0; q < 6; q++) sf[(int64_t)q * A.ns + m] = s[q];
    if (vm)
        vm[m] = sqrt(0.5 * ((s[0] - s[1]) * (s[0] - s[1]) + (s[1] - s[2]) * (s[1] - s[2]) +
                            (s[2] - s[0]) * (s[2] - s[0]) + 6.0 * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5])));
}

__global__ __launch_bounds__(kBlock) void k_assig(int64_t ns, const uint16_t *pix, unsigned *counts)
{
    const int64_t m = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (m < ns) atomicAdd(&counts[pix[m]], 1u);
}

__global__ __launch_bounds__(kBlock) void k_to_f32(const double *s, float *d, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) d[i] = (float)s[i];
}

__global__ __launch_bounds__(kBlock) void k_to_f64(const float *s, double *d, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) d[i] = (double)s[i];
}

// ---------------------------------------------------------------------------
JumpArgs jump_args(const Solver &s)
{
    JumpArgs A{};
    A.nx = s.nx; A.ny = s.ny; A.nz = s.nz; A.nrhs = s.nrhs; A.ns = s.ns;
    A.pix = s.pix; A.dk = s.dk;
    for (int q = 0; q < 6 * kMaxRhs; q++) A.e[q] = s.strain[q];
    return A;
}

PostArgs post_args(const Solver &s)
{
    PostArgs A{};
    A.nx = s.nx; A.ny = s.ny; A.nz = s.nz; A.nrhs = s.nrhs; A.ns = s.ns;
    A.pix = s.pix; A.cmod = s.cmod; A.u = s.u;
    for (int q = 0; q < 6 * kMaxRhs; q++) A.e[q] = s.strain[q];
    return A;
}

template <typename T>
StencilArgs<T> stencil_args(const Solver &s, const T *dk, const T *sten, const T *x, T *y)
{
    StencilArgs<T> A{};
    A.nx = s.nx; A.ny = s.ny; A.nz = s.nz;
    A.ntx = (s.nx + TX - 1) / TX; A.nty = (s.ny + TY - 1) / TY;
    A.ns = s.ns; A.pix = s.pix; A.dk = dk; A.sten = sten; A.x = x; A.y = y;
    A.partials = s.partials; A.counter = s.counter; A.st = s.st;
    A.results = s.results; A.cconst = s.cconst;
    return A;
}

template <typename T, int MODE>
void dispatch_stencil(const Solver &s, const StencilArgs<T> &A)
{
    const dim3 grid((unsigned)stencil_blocks(s)), block(kBlock);
    switch (s.nrhs) {
    case 1: hipLaunchKernelGGL((k_stencil<T, 1, MODE>), grid, block, 0, s.stream, A); break;
    case 2: hipLaunchKernelGGL((k_stencil<T, 2, MODE>), grid, block, 0, s.stream, A); break;
    case 3: hipLaunchKernelGGL((k_stencil<T, 3, MODE>), grid, block, 0, s.stream, A); break;
    case 4: hipLaunchKernelGGL((k_stencil<T, 4, MODE>), grid, block, 0, s.stream, A); break;
    case 5: hipLaunchKernelGGL((k_stencil<T, 5, MODE>), grid, block, 0, s.stream, A); break;
    default: hipLaunchKernelGGL((k_stencil<T, 6, MODE>), grid, block, 0, s.stream, A); break;
    }
}

template <typename T>
PcgArgs<T> pcg_args(const Solver &s, int it)
{
    PcgArgs<T> A{};
    A.ns = s.ns; A.it = it; A.u = s.u;
    A.partials = s.partials; A.counter = s.counter; A.st = s.st; A.hist = s.hist;
    if constexpr (sizeof(T) == 4) {
        A.g = s.g32; A.h = s.h32; A.z = s.z32; A.ah = s.ah32; A.minv = s.minv32;
    } else {
        A.g = s.g; A.h = s.h; A.z = s.z; A.ah = s.ah; A.minv = s.minv;
    }
    return A;
}

#define E3X_DISPATCH_NR(KERNEL, T, grid, block, stream, args, ...)                              \
    switch (s.nrhs) {                                                                           \
    case 1: hipLaunchKernelGGL((KERNEL<T, 1, ##__VA_ARGS__>), grid, block, 0, stream, args); break; \
    case 2: hipLaunchKernelGGL((KERNEL<T, 2, ##__VA_ARGS__>), grid, block, 0, stream, args); break; \
    case 3: hipLaunchKernelGGL((KERNEL<T, 3, ##__VA_ARGS__>), grid, block, 0, stream, args); break; \
    case 4: hipLaunchKernelGGL((KERNEL<T, 4, ##__VA_ARGS__>), grid, block, 0, stream, args); break; \
    case 5: hipLaunchKernelGGL((KERNEL<T, 5, ##__VA_ARGS__>), grid, block, 0, stream, args); break; \
    default: hipLaunchKernelGGL((KERNEL<T, 6, ##__VA_ARGS__>), grid, block, 0, stream, args); break; \
    }

template <typename T>
void pcg_iter_t(Solver &s, int it, bool replace)
{
    const dim3 vg((unsigned)vector_blocks(s)), blk(kBlock);
    // 1. Ah = A h, hAh, lambda (f90:1578-1907)
    StencilArgs<T> SA;
    if constexpr (sizeof(T) == 4) SA = stencil_args<T>(s, s.dk32, s.sten32, s.h32, s.ah32);
    else SA = stencil_args<T>(s, s.dk, s.sten, s.h, s.ah);
    SA.it = it;
    dispatch_stencil<T, kApply>(s, SA);
    PcgArgs<T> PA = pcg_args<T>(s, it);
    if (!replace) {
        // 2. u, g update, z = M^-1 g, z.g and the scalar bookkeeping (f90:1914-2007)
        E3X_DISPATCH_NR(k_pcg_update, T, vg, blk, s.stream, PA, true, true, true);
    } else {
        // 2'. u update only, true gradient g = K u + b in FP64, then z and z.g
        E3X_DISPATCH_NR(k_pcg_update, T, vg, blk, s.stream, PA, false, true, false);
        StencilArgs<double> RA = stencil_args<double>(s, s.dk, s.sten, s.u, nullptr);
        RA.b = s.b; RA.y32 = s.g32; RA.it = it;
        dispatch_stencil<double, kResid>(s, RA);
        E3X_DISPATCH_NR(k_pcg_update, T, vg, blk, s.stream, PA, false, false, true);
    }
    // 3. h = z (+ gamma h)
    E3X_DISPATCH_NR(k_pcg_direction, T, vg, blk, s.stream, PA);
}

}  // namespace

int64_t stencil_blocks(const Solver &s)
{
    return (int64_t)((s.nx + TX - 1) / TX) * ((s.ny + TY - 1) / TY) * ((s.nz + TZ - 1) / TZ);
}

int64_t vector_blocks(const Solver &s) { return (s.ns + kBlock - 1) / kBlock; }

hipError_t launch_femat(Solver &s, double c11, double c12, double c44, const double *pm)
{
    hipLaunchKernelGGL(k_femat, dim3(s.nphase), dim3(kBlock), 0, s.stream, s.nphase, s.flag_m,
                       s.orient, c11, c12, c44, pm[0], pm[1], pm[2], pm[3], s.cmod, s.dk, s.sten,
                       s.dk32, s.sten32);
    return hipGetLastError();
}

hipError_t launch_rhs(Solver &s)
{
    hipLaunchKernelGGL(k_rhs, dim3((unsigned)vector_blocks(s)), dim3(kBlock), 0, s.stream, jump_args(s),
                       s.b, s.partials, s.counter, s.cconst);
    return hipGetLastError();
}

hipError_t launch_affine(Solver &s)
{
    hipLaunchKernelGGL(k_affine, dim3((unsigned)vector_blocks(s)), dim3(kBlock), 0, s.stream, s.nx, s.ny,
                       s.nz, s.nrhs, s.ns, jump_args(s), s.u);
    return hipGetLastError();
}

hipError_t launch_energy(Solver &s)
{
    StencilArgs<double> A = stencil_args<double>(s, s.dk, s.sten, s.u, s.g);
    A.b = s.b;
    dispatch_stencil<double, kEnergy>(s, A);
    return hipGetLastError();
}

hipError_t launch_apply_plain(Solver &s, const double *x, double *y)
{
    StencilArgs<double> A = stencil_args<double>(s, s.dk, s.sten, x, y);
    dispatch_stencil<double, kPlain>(s, A);
    return hipGetLastError();
}

hipError_t launch_precond(Solver &s, bool want32)
{
    hipLaunchKernelGGL(k_precond, dim3((unsigned)vector_blocks(s)), dim3(kBlock), 0, s.stream, s.nx, s.ny,
                       s.nz, s.ns, s.pix, s.dk, s.minv, want32 ? s.minv32 : nullptr);
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
    }
    hipError_t e = hipMemcpyAsync(s.st, hs, sizeof(PcgState) * s.nrhs, hipMemcpyHostToDevice, s.stream);
    if (e != hipSuccess) return e;
    const dim3 vg((unsigned)vector_blocks(s)), blk(kBlock);
    if (mixed) {
        PcgArgs<float> A = pcg_args<float>(s, 0);
        A.kkk = kkk; A.ldemb = ldemb;
        E3X_DISPATCH_NR(k_pcg_init, float, vg, blk, s.stream, A);
    } else {
        PcgArgs<double> A = pcg_args<double>(s, 0);
        A.kkk = kkk; A.ldemb = ldemb;
        E3X_DISPATCH_NR(k_pcg_init, double, vg, blk, s.stream, A);
    }
    return hipGetLastError();
}

hipError_t launch_pcg_iter(Solver &s, bool mixed, int it, bool replace)
{
    if (mixed) pcg_iter_t<float>(s, it, replace);
    else pcg_iter_t<double>(s, it, false);
    return hipGetLastError();
}

hipError_t launch_stress(Solver &s)
{
    const dim3 vg((unsigned)vector_blocks(s)), blk(kBlock);
    const PostArgs A = post_args(s);
    switch (s.nrhs) {
    case 1: hipLaunchKernelGGL(k_stress<1>, vg, blk, 0, s.stream, A, s.partials, s.counter, s.results); break;
    case 2: hipLaunchKernelGGL(k_stress<2>, vg, blk, 0, s.stream, A, s.partials, s.counter, s.results); break;
    case 3: hipLaunchKernelGGL(k_stress<3>, vg, blk, 0, s.stream, A, s.partials, s.counter, s.results); break;
    case 4: hipLaunchKernelGGL(k_stress<4>, vg, blk, 0, s.stream, A, s.partials, s.counter, s.results); break;
    case 5: hipLaunchKernelGGL(k_stress<5>, vg, blk, 0, s.stream, A, s.partials, s.counter, s.results); break;
    default: hipLaunchKernelGGL(k_stress<6>, vg, blk, 0, s.stream, A, s.partials, s.counter, s.results); break;
    }
    return hipGetLastError();
}

hipError_t launch_fullfield(Solver &s, int rhs, double *sf, double *vm)
{
    hipLaunchKernelGGL(k_fullfield, dim3((unsigned)vector_blocks(s)), dim3(kBlock), 0, s.stream,
                       post_args(s), rhs, sf, vm);
    return hipGetLastError();
}

hipError_t launch_assig(Solver &s)
{
    hipError_t e = hipMemsetAsync(s.counts, 0, sizeof(unsigned) * s.nphase, s.stream);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_assig, dim3((unsigned)vector_blocks(s)), dim3(kBlock), 0, s.stream, s.ns, s.pix,
                       s.counts);
    return hipGetLastError();
}

hipError_t launch_to_f32(const double *src, float *dst, int64_t n, hipStream_t st)
{
    hipLaunchKernelGGL(k_to_f32, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, src, dst, n);
    return hipGetLastError();
}

hipError_t launch_to_f64(const float *src, double *dst, int64_t n, hipStream_t st)
{
    hipLaunchKernelGGL(k_to_f64, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, src, dst, n);
    return hipGetLastError();
}

}  // namespace e3x
