// gfx950 post-processing kernels of the elas3dxtal_pcg solve: volume-averaged and
// full-field voxel-centre stress/strain, and precision conversions.
#include "e3x_device.hpp"

#include <cmath>
#include <mutex>
#include <string>

namespace e3x {
namespace {

// ---------------------------------------------------------------------------
// Voxel-centre strain and stress (stress_OpenMP f90:2149-2408,
// stress_fullfield_OpenMP f90:2639-2833).  es at the centre is +-1/4.
struct PostArgs {
    int nx, ny, nz, nrhs;
    int64_t ns;
    const void *pix;         // 0-based phase ids (a handle's copy), uint16 or (wide) uint32
    int wide;
    int64_t mstride;         // batch: nodes between the members' phase maps (0: one shared map)
    const int32_t *pix32;    // DEV form: the caller's labels 1..nphase
    int nphase;              // DEV form: labels outside [1, nphase] give NaN
    const double *cmod, *u;  // u: [nrhs][ns][3], DEV form [nrhs][3][ns] (u(ns,3) per rhs)
    double e[6 * kMaxRhs];
};

// DEV: the caller's arrays as the reference lays them out (int32 labels, u(ns,3)).
template <bool DEV>
__device__ void voxel_strain_stress(const PostArgs &A, int r, int64_t m, double eps[6], double sig[6])
{
    const int i = (int)(m % A.nx), j = (int)((m / A.nx) % A.ny), k = (int)(m / ((int64_t)A.nx * A.ny));
    const double *u = A.u + (int64_t)r * 3 * A.ns;
    const double *e = A.e + 6 * r;  // exx eyy ezz exz eyz exy
    double uu[8][3];
    for (int a = 0; a < 8; a++) {
        const int64_t q = wrap(i + PX(a), A.nx) + (int64_t)A.nx * (wrap(j + PY(a), A.ny) + (int64_t)A.ny * wrap(k + PZ(a), A.nz));
        for (int c = 0; c < 3; c++) uu[a][c] = DEV ? u[c * A.ns + q] : u[3 * q + c];
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
    int p;
    if constexpr (DEV) {
        p = A.pix32[m] - 1;
        if (p < 0 || p >= A.nphase) {  // an unchecked caller label: no table read, NaN out
            for (int q = 0; q < 6; q++) sig[q] = __builtin_nan("");
            return;
        }
    } else {
        p = phase_at(A.pix, A.wide, r * A.mstride + m);
    }
    const double *c = A.cmod + (int64_t)p * 36;
    for (int q = 0; q < 6; q++) {
        double s = 0.0;
        for (int t = 0; t < 6; t++) s += c[q * 6 + t] * eps[t];
        sig[q] = s;
    }
}

template <int NR, bool DEV = false>
__global__ __launch_bounds__(kBlock) void k_stress(PostArgs A, double *partials)
{
    double part[12 * NR];
#pragma unroll
    for (int q = 0; q < 12 * NR; q++) part[q] = 0.0;
    for (int64_t m = (int64_t)blockIdx.x * kBlock + threadIdx.x; m < A.ns; m += (int64_t)gridDim.x * kBlock) {
#pragma unroll
        for (int r = 0; r < NR; r++) {
            double eps[6], sig[6];
            voxel_strain_stress<DEV>(A, r, m, eps, sig);
#pragma unroll
            for (int q = 0; q < 6; q++) {
                part[12 * r + q] += sig[q];
                part[12 * r + 6 + q] += eps[q];
            }
        }
    }
    write_partials<12 * NR>(part, partials);
}

template <int NR>
__global__ __launch_bounds__(kFinBlock) void k_stress_finish(const double *partials, int nparts, int64_t ns,
                                                          double *results)
{
    double tot[12 * NR];
    sum_partials<12 * NR>(partials, nparts, tot);
    if (threadIdx.x == 0)
        for (int q = 0; q < 12 * NR; q++) results[q] = tot[q] / (double)ns;
}

__global__ __launch_bounds__(kBlock) void k_fullfield(PostArgs A, int r, double *sf, double *vm)
{
    const int64_t m = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (m >= A.ns) return;
    double eps[6], s[6];
    voxel_strain_stress<false>(A, r, m, eps, s);
    if (sf)
        for (int q = 0; q < 6; q++) sf[(int64_t)q * A.ns + m] = s[q];
    if (vm)
        vm[m] = sqrt(0.5 * ((s[0] - s[1]) * (s[0] - s[1]) + (s[1] - s[2]) * (s[1] - s[2]) +
                            (s[2] - s[0]) * (s[2] - s[0]) + 6.0 * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5])));
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

// Between the reference's field layout u(ns, 3) (component planes) and the device's
// node-interleaved [ns][3]; one thread per node.
__global__ __launch_bounds__(kBlock) void k_planes_to_nodes(const double *s, double *d, int64_t ns)
{
    const int64_t m = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (m < ns) {
        const double a = s[m], b = s[ns + m], c = s[2 * ns + m];
        d[3 * m] = a; d[3 * m + 1] = b; d[3 * m + 2] = c;
    }
}

__global__ __launch_bounds__(kBlock) void k_nodes_to_planes(const double *s, double *d, int64_t ns)
{
    const int64_t m = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (m < ns) {
        const double a = s[3 * m], b = s[3 * m + 1], c = s[3 * m + 2];
        d[m] = a; d[ns + m] = b; d[2 * ns + m] = c;
    }
}

PostArgs post_args(const Solver &s)
{
    PostArgs A{};
    A.nx = s.nx; A.ny = s.ny; A.nz = s.nz; A.nrhs = s.nrhs; A.ns = s.ns;
    A.pix = s.pix; A.wide = s.wide; A.mstride = s.mstride(); A.cmod = s.cmod; A.u = s.u;
    for (int q = 0; q < 6 * kMaxRhs; q++) A.e[q] = s.strain[q];
    return A;
}

template <int NR>
void stress_nr(Solver &s)
{
    const int g = (int)vector_grid(s);
    hipLaunchKernelGGL(k_stress<NR>, dim3(g), dim3(kBlock), 0, s.stream, post_args(s), s.partials);
    hipLaunchKernelGGL(k_stress_finish<NR>, dim3(1), dim3(kFinBlock), 0, s.stream, (const double *)s.partials, g, s.ns,
                       s.results);
}

// e3x_stress_dev: the averages of caller-owned device arrays, with the block partials in
// this file's scratch array (the grid of launch_stress, so the sums run in the same order).
template <int NR>
void stress_dev_nr(const PostArgs &A, int g, double *scratch, double *out, hipStream_t st)
{
    hipLaunchKernelGGL((k_stress<NR, true>), dim3(g), dim3(kBlock), 0, st, A, scratch);
    hipLaunchKernelGGL(k_stress_finish<NR>, dim3(1), dim3(kFinBlock), 0, st, (const double *)scratch, g, A.ns, out);
}

}  // namespace

// ---------------------------------------------------------------------------
// Scratch of e3x_stress_dev: the block partials of the averages, 12 per right-hand side
// per block of the streaming grid (at most kVecGridMax blocks), in device memory the
// library holds for the life of the process on every GPU it runs on (1.2 MB).  Host code
// reaches a device's copy only through the address hipGetSymbolAddress returns for that
// device, never through the array's name.
constexpr int64_t kStressScratch = 12 * (int64_t)kMaxRhs * kVecGridMax;
static __device__ __align__(256) double scratch[kStressScratch];

namespace {
constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kScratchDevices] = {};

// This device's copy of the scratch array (looked up once per device, with it current).
hipError_t scratch_for(int dev, double **out)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}
}  // namespace

int stress_dev(int nx, int ny, int nz, int nphase, int nrhs, const int32_t *pix, const double *cmod,
               const double *strain6, const double *u, double *averages, hipStream_t st)
{
    if (nx < 2 || ny < 2 || nz < 2) return report_error(1, "grid extents must be >= 2");
    if ((int64_t)nx * ny * nz > (int64_t)INT32_MAX) return report_error(1, "grid exceeds 2^31 nodes");
    if (nphase < 1) return report_error(1, "nphase must be >= 1");
    if (nrhs < 1 || nrhs > kMaxRhs)
        return report_error(1, "nrhs must be in [1, 6]: the scratch holds the partials of at most 6 right-hand sides");
    if (!pix || !cmod || !strain6 || !u || !averages) return report_error(1, "NULL argument");
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kScratchDevices)
        return report_error(2, "no current HIP device (or device index beyond 63)");
    double *sc = nullptr;
    if (hipError_t e = scratch_for(dev, &sc); e != hipSuccess)
        return report_error(2, (std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e)).c_str());
    PostArgs A{};
    A.nx = nx; A.ny = ny; A.nz = nz; A.nrhs = nrhs; A.ns = (int64_t)nx * ny * nz;
    A.pix32 = pix; A.nphase = nphase; A.cmod = cmod; A.u = u;
    for (int q = 0; q < 6 * nrhs; q++) A.e[q] = strain6[q];
    const int64_t vb = (A.ns + kBlock - 1) / kBlock;
    const int g = (int)(vb < kVecGridMax ? vb : kVecGridMax);  // vector_grid(): launch_stress's grid
    static_assert(kStressScratch >= 12 * (int64_t)kMaxRhs * kVecGridMax, "scratch too small");
    switch (nrhs) {
    case 1: stress_dev_nr<1>(A, g, sc, averages, st); break;
    case 2: stress_dev_nr<2>(A, g, sc, averages, st); break;
    case 3: stress_dev_nr<3>(A, g, sc, averages, st); break;
    case 4: stress_dev_nr<4>(A, g, sc, averages, st); break;
    case 5: stress_dev_nr<5>(A, g, sc, averages, st); break;
    default: stress_dev_nr<6>(A, g, sc, averages, st); break;
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return report_error(2, (std::string("stress launch: ") + hipGetErrorString(e)).c_str());
    return 0;
}

hipError_t launch_stress(Solver &s)
{
    switch (s.nrhs) {
    case 1: stress_nr<1>(s); break;
    case 2: stress_nr<2>(s); break;
    case 3: stress_nr<3>(s); break;
    case 4: stress_nr<4>(s); break;
    case 5: stress_nr<5>(s); break;
    default: stress_nr<6>(s); break;
    }
    return hipGetLastError();
}

hipError_t launch_fullfield(Solver &s, int rhs, double *sf, double *vm)
{
    hipLaunchKernelGGL(k_fullfield, dim3((unsigned)vector_blocks(s)), dim3(kBlock), 0, s.stream,
                       post_args(s), rhs, sf, vm);
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

hipError_t launch_planes_to_nodes(const double *src, double *dst, int64_t ns, hipStream_t st)
{
    hipLaunchKernelGGL(k_planes_to_nodes, dim3((unsigned)((ns + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, src,
                       dst, ns);
    return hipGetLastError();
}

hipError_t launch_nodes_to_planes(const double *src, double *dst, int64_t ns, hipStream_t st)
{
    hipLaunchKernelGGL(k_nodes_to_planes, dim3((unsigned)((ns + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, src,
                       dst, ns);
    return hipGetLastError();
}

}  // namespace e3x
