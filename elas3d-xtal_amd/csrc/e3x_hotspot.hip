// gfx950 hotspot pass over the von Mises field: the steps the reference's MATLAB
// post-processing takes for top_peaks.txt (applications/sec3_2/A_2_0/
// elas3dxtal_postprocessing.m:86-243, the same in the sec3_3 scripts).
//
//   k_hot_smooth   masked normalised Gaussian smoothing, zero padding (.m:86-97)
//   k_hot_regmax   regional-maximum test against the 26 neighbours, threshold (.m:206-208);
//                  voxels with an equal neighbour go to a second list, whose equal-valued
//                  components the host completes (plateaus are rare)
//   k_hot_topk     the first npeaks candidates by value, descending, ties by index (.m:215)
//
// Each voxel is one lane; the field is read through L2 (27 values per lane, 14 B of HBM
// traffic per voxel for the smoothing, 4 B for the test).
#include "e3x_device.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace e3x {
namespace {

constexpr int kHotRadiusMax = 3;  // sigma <= 1.5 (radius ceil(2 sigma))

struct HotArgs {
    int nx, ny, nz, nsolid, r;
    int64_t ns;
    const void *pix;  // 0-based phase id per voxel, uint16 or (wide) uint32
    int wide;
    const double *vm;
    float *sm;            // smoothed field
    double w[kHotRadiusMax + 1];   // normalised 1-D weights at |d| = 0..r
    float wf[kHotRadiusMax + 1];
    float thr;            // MPa
    HotEntry *cand, *plat;
    unsigned long long *count;  // [0] strict maxima, [1] voxels with an equal neighbour
};

__device__ __forceinline__ bool solid(const HotArgs &A, int64_t m) { return phase_at(A.pix, A.wide, m) + 1 <= A.nsolid; }

// s = G*(vm . solid) / G*(solid) on solid voxels, 0 elsewhere (.m:86-97).  The numerator
// is summed in double, the denominator in single, and the quotient taken in single, as
// the script's double ./ single gives; a denominator below 1e-6 counts as 1 (.m:93).
__global__ __launch_bounds__(kBlock) void k_hot_smooth(HotArgs A)
{
    const int64_t m = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (m >= A.ns) return;
    if (!solid(A, m)) {
        A.sm[m] = 0.0f;
        return;
    }
    const int nx = A.nx, ny = A.ny, nz = A.nz, r = A.r;
    const int i = (int)(m % nx), j = (int)((m / nx) % ny), k = (int)(m / ((int64_t)nx * ny));
    double num = 0.0;
    float den = 0.0f;
    for (int dz = -r; dz <= r; dz++) {
        const int kk = k + dz;
        if (kk < 0 || kk >= nz) continue;  // zero padding: the grid is not wrapped here
        const double wz = A.w[dz < 0 ? -dz : dz];
        const float wzf = A.wf[dz < 0 ? -dz : dz];
        for (int dy = -r; dy <= r; dy++) {
            const int jj = j + dy;
            if (jj < 0 || jj >= ny) continue;
            const double wzy = wz * A.w[dy < 0 ? -dy : dy];
            const float wzyf = wzf * A.wf[dy < 0 ? -dy : dy];
            const int64_t row = ((int64_t)kk * ny + jj) * nx;
            for (int dx = -r; dx <= r; dx++) {
                const int ii = i + dx;
                if (ii < 0 || ii >= nx) continue;
                const int64_t q = row + ii;
                if (!solid(A, q)) continue;
                num += wzy * A.w[dx < 0 ? -dx : dx] * A.vm[q];
                den += wzyf * A.wf[dx < 0 ? -dx : dx];
            }
        }
    }
    if (den < 1e-6f) den = 1.0f;
    A.sm[m] = (float)num / den;
}

// Voxels at or above the threshold that no in-grid neighbour exceeds; with no equal
// neighbour they are regional maxima by themselves, otherwise the host decides on their
// equal-valued component (.m:206-208: imregionalmax, 26-connectivity, no wrap).
__global__ __launch_bounds__(kBlock) void k_hot_regmax(HotArgs A)
{
    const int64_t m = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (m >= A.ns) return;
    const float v = A.sm[m];
    if (!(v / 1e6f >= A.thr)) return;
    const int nx = A.nx, ny = A.ny, nz = A.nz;
    const int i = (int)(m % nx), j = (int)((m / nx) % ny), k = (int)(m / ((int64_t)nx * ny));
    bool eq = false;
    for (int dz = -1; dz <= 1; dz++) {
        const int kk = k + dz;
        if (kk < 0 || kk >= nz) continue;
        for (int dy = -1; dy <= 1; dy++) {
            const int jj = j + dy;
            if (jj < 0 || jj >= ny) continue;
            for (int dx = -1; dx <= 1; dx++) {
                const int ii = i + dx;
                if (ii < 0 || ii >= nx || (dx == 0 && dy == 0 && dz == 0)) continue;
                const float u = A.sm[((int64_t)kk * ny + jj) * nx + ii];
                if (u > v) return;
                eq |= u == v;
            }
        }
    }
    const unsigned long long slot = atomicAdd(&A.count[eq ? 1 : 0], 1ull);
    HotEntry *list = eq ? A.plat : A.cand;
    list[slot] = HotEntry{v, 0, m};  // at most ns entries per list, each list holds ns
}

// a before b: larger value first, then smaller index
__device__ __forceinline__ bool hot_before(float av, int64_t am, float bv, int64_t bm)
{
    return av > bv || (av == bv && am < bm);
}

// One block: npeaks passes, each the first entry after the previous pick in the order.
__global__ __launch_bounds__(kBlock) void k_hot_topk(const HotEntry *cand, const unsigned long long *count, int npeaks,
                                                    HotEntry *out)
{
    __shared__ float bv[kBlock];
    __shared__ int64_t bm[kBlock];
    const int64_t n = (int64_t)count[0];
    const int t = threadIdx.x;
    float pv = INFINITY;
    int64_t pm = -1;
    for (int p = 0; p < npeaks; p++) {
        float best_v = -INFINITY;
        int64_t best_m = -1;
        for (int64_t e = t; e < n; e += kBlock) {
            const float v = cand[e].v;
            const int64_t m = cand[e].m;
            if (!hot_before(pv, pm, v, m)) continue;  // taken already
            if (best_m < 0 || hot_before(v, m, best_v, best_m)) { best_v = v; best_m = m; }
        }
        bv[t] = best_v;
        bm[t] = best_m;
        __syncthreads();
        for (int s = kBlock / 2; s > 0; s >>= 1) {
            if (t < s) {
                const float ov = bv[t + s];
                const int64_t om = bm[t + s];
                if (om >= 0 && (bm[t] < 0 || hot_before(ov, om, bv[t], bm[t]))) { bv[t] = ov; bm[t] = om; }
            }
            __syncthreads();
        }
        pv = bv[0];
        pm = bm[0];
        __syncthreads();
        if (pm < 0) break;  // no entry left (uniform over the block); the host reads count[0]
        if (t == 0) out[p] = HotEntry{pv, 0, pm};
    }
}

}  // namespace

hipError_t launch_hotspots(Solver &s, int rhs, int nsolid, double sigma, float thr_mpa, int npeaks, HotResult &res)
{
    const int64_t ns = s.ns;
    HotArgs A{};
    A.nx = s.nx; A.ny = s.ny; A.nz = s.nz; A.ns = ns; A.nsolid = nsolid;
    // the phase map of the right-hand side's member (a batch), else the shared one
    A.pix = (const char *)s.pix + (s.wide ? 4 : 2) * (int64_t)rhs * s.mstride();
    A.wide = s.wide ? 1 : 0;
    A.vm = s.post + 6 * ns;
    A.sm = reinterpret_cast<float *>(s.post);
    A.cand = reinterpret_cast<HotEntry *>(s.post + ns);       // planes 1-2: ns entries of 16 B
    A.plat = reinterpret_cast<HotEntry *>(s.post + 3 * ns);   // planes 3-4
    A.count = reinterpret_cast<unsigned long long *>(s.hot);
    HotEntry *top = reinterpret_cast<HotEntry *>(s.hot + 2);
    // imgaussfilt3's kernel: exp(-d^2 / (2 sigma^2)) over |d| <= ceil(2 sigma), normalised
    A.r = (int)std::ceil(2.0 * sigma);
    double tot = 0.0;
    for (int d = -A.r; d <= A.r; d++) tot += std::exp(-(double)d * d / (2.0 * sigma * sigma));
    for (int d = 0; d <= A.r; d++) {
        A.w[d] = std::exp(-(double)d * d / (2.0 * sigma * sigma)) / tot;
        A.wf[d] = (float)A.w[d];
    }
    A.thr = thr_mpa;
    const dim3 grid((unsigned)((ns + kBlock - 1) / kBlock)), blk(kBlock);
    hipError_t e = hipMemsetAsync(s.hot, 0, 2 * sizeof(unsigned long long), s.stream);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_hot_smooth, grid, blk, 0, s.stream, A);
    hipLaunchKernelGGL(k_hot_regmax, grid, blk, 0, s.stream, A);
    hipLaunchKernelGGL(k_hot_topk, dim3(1), blk, 0, s.stream, (const HotEntry *)A.cand,
                       (const unsigned long long *)A.count, npeaks, top);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    unsigned long long cnt[2];
    std::vector<HotEntry> best((size_t)npeaks);
    if ((e = hipMemcpyAsync(cnt, A.count, sizeof(cnt), hipMemcpyDeviceToHost, s.stream)) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(best.data(), top, sizeof(HotEntry) * npeaks, hipMemcpyDeviceToHost, s.stream)) != hipSuccess)
        return e;
    if ((e = hipStreamSynchronize(s.stream)) != hipSuccess) return e;
    best.resize((size_t)std::min<unsigned long long>(cnt[0], (unsigned long long)npeaks));
    int64_t found = (int64_t)cnt[0];

    if (cnt[1] > 0) {
        // Voxels with an equal neighbour: a component of equal values is a regional maximum
        // when no neighbour of any member is larger; then every member is one.
        std::vector<float> sm((size_t)ns);
        std::vector<HotEntry> plat((size_t)cnt[1]);
        if ((e = hipMemcpyAsync(sm.data(), A.sm, sizeof(float) * ns, hipMemcpyDeviceToHost, s.stream)) != hipSuccess)
            return e;
        if ((e = hipMemcpyAsync(plat.data(), A.plat, sizeof(HotEntry) * cnt[1], hipMemcpyDeviceToHost, s.stream)) !=
            hipSuccess)
            return e;
        if ((e = hipStreamSynchronize(s.stream)) != hipSuccess) return e;
        std::vector<uint8_t> seen((size_t)ns, 0);
        std::vector<int64_t> comp, stack;
        const int nx = s.nx, ny = s.ny, nz = s.nz;
        for (const HotEntry &p : plat) {
            if (seen[(size_t)p.m]) continue;
            const float v = p.v;
            bool top_ok = true;
            comp.clear();
            stack.assign(1, p.m);
            seen[(size_t)p.m] = 1;
            while (!stack.empty()) {
                const int64_t m = stack.back();
                stack.pop_back();
                comp.push_back(m);
                const int i = (int)(m % nx), j = (int)((m / nx) % ny), k = (int)(m / ((int64_t)nx * ny));
                for (int dz = -1; dz <= 1; dz++)
                    for (int dy = -1; dy <= 1; dy++)
                        for (int dx = -1; dx <= 1; dx++) {
                            const int ii = i + dx, jj = j + dy, kk = k + dz;
                            if (ii < 0 || ii >= nx || jj < 0 || jj >= ny || kk < 0 || kk >= nz) continue;
                            const int64_t q = ((int64_t)kk * ny + jj) * nx + ii;
                            const float u = sm[(size_t)q];
                            if (u > v) top_ok = false;
                            if (u == v && !seen[(size_t)q]) {
                                seen[(size_t)q] = 1;
                                stack.push_back(q);
                            }
                        }
            }
            if (!top_ok) continue;
            found += (int64_t)comp.size();
            for (int64_t m : comp) best.push_back(HotEntry{v, 0, m});
        }
        std::sort(best.begin(), best.end(), [](const HotEntry &a, const HotEntry &b) {
            return a.v > b.v || (a.v == b.v && a.m < b.m);
        });
        if ((int64_t)best.size() > npeaks) best.resize((size_t)npeaks);
    }
    res.found = found;
    res.peaks = best;
    return hipSuccess;
}

}  // namespace e3x
