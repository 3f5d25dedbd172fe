// Host-side I/O of the elas3dxtal_pcg solver: HDF5 in the reference's schema and a
// seeded periodic microstructure generator (include/e3x_io.h).
#include "e3x_io.h"

#include <dlfcn.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <random>
#include <string>
#include <vector>

#include "e3x.h"
#include "e3x_serve.h"

#include <cstdlib>
#include <map>

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

// ------------------------------------------------------------------ HDF5 1.10 C API
// Loaded with dlopen so the library carries no link-time dependency on the HDF5 build
// (and none on the C++ runtime that ships next to it).
typedef int64_t hid_t;
typedef int herr_t;
typedef int htri_t;
typedef unsigned long long hsize_t;
constexpr hid_t kP_DEFAULT = 0, kS_ALL = 0, kE_DEFAULT = 0;
constexpr unsigned kF_ACC_RDONLY = 0x0000u, kF_ACC_TRUNC = 0x0002u;

struct H5 {
    void *so = nullptr;
    herr_t (*open)(void);
    herr_t (*eset_auto)(hid_t, void *, void *);
    hid_t (*fopen)(const char *, unsigned, hid_t);
    hid_t (*fcreate)(const char *, unsigned, hid_t, hid_t);
    herr_t (*fclose)(hid_t);
    hid_t (*dopen)(hid_t, const char *, hid_t);
    hid_t (*dcreate)(hid_t, const char *, hid_t, hid_t, hid_t, hid_t, hid_t);
    herr_t (*dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void *);
    herr_t (*dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void *);
    herr_t (*dclose)(hid_t);
    hid_t (*dget_space)(hid_t);
    hid_t (*screate_simple)(int, const hsize_t *, const hsize_t *);
    herr_t (*sclose)(hid_t);
    int (*sndims)(hid_t);
    int (*sdims)(hid_t, hsize_t *, hsize_t *);
    htri_t (*aexists)(hid_t, const char *);
    hid_t (*aopen)(hid_t, const char *, hid_t);
    hid_t (*acreate)(hid_t, const char *, hid_t, hid_t, hid_t, hid_t);
    herr_t (*aread)(hid_t, hid_t, void *);
    herr_t (*awrite)(hid_t, hid_t, const void *);
    herr_t (*aclose)(hid_t);
    hid_t aget_space_dummy;
    hid_t *t_int32 = nullptr, *t_double = nullptr;
    bool ok = false;
    std::string why;
};

H5 &h5()
{
    static H5 H;
    static std::once_flag once;
    std::call_once(once, [] {
        const char *env = getenv("E3X_HDF5_LIB");
        const char *cands[] = {env, "libhdf5.so.103", "/opt/conda/lib/libhdf5.so.103", "libhdf5.so"};
        for (const char *c : cands) {
            if (!c) continue;
            H.so = dlopen(c, RTLD_NOW | RTLD_LOCAL);
            if (H.so) break;
        }
        if (!H.so) {
            H.why = "HDF5 1.10 shared library (libhdf5.so.103) not found; set E3X_HDF5_LIB";
            return;
        }
        bool good = true;
        auto sym = [&](const char *n) {
            void *p = dlsym(H.so, n);
            if (!p) good = false;
            return p;
        };
        H.open = (herr_t(*)(void))sym("H5open");
        H.eset_auto = (herr_t(*)(hid_t, void *, void *))sym("H5Eset_auto2");
        H.fopen = (hid_t(*)(const char *, unsigned, hid_t))sym("H5Fopen");
        H.fcreate = (hid_t(*)(const char *, unsigned, hid_t, hid_t))sym("H5Fcreate");
        H.fclose = (herr_t(*)(hid_t))sym("H5Fclose");
        H.dopen = (hid_t(*)(hid_t, const char *, hid_t))sym("H5Dopen2");
        H.dcreate = (hid_t(*)(hid_t, const char *, hid_t, hid_t, hid_t, hid_t, hid_t))sym("H5Dcreate2");
        H.dread = (herr_t(*)(hid_t, hid_t, hid_t, hid_t, hid_t, void *))sym("H5Dread");
        H.dwrite = (herr_t(*)(hid_t, hid_t, hid_t, hid_t, hid_t, const void *))sym("H5Dwrite");
        H.dclose = (herr_t(*)(hid_t))sym("H5Dclose");
        H.dget_space = (hid_t(*)(hid_t))sym("H5Dget_space");
        H.screate_simple = (hid_t(*)(int, const hsize_t *, const hsize_t *))sym("H5Screate_simple");
        H.sclose = (herr_t(*)(hid_t))sym("H5Sclose");
        H.sndims = (int (*)(hid_t))sym("H5Sget_simple_extent_ndims");
        H.sdims = (int (*)(hid_t, hsize_t *, hsize_t *))sym("H5Sget_simple_extent_dims");
        H.aexists = (htri_t(*)(hid_t, const char *))sym("H5Aexists");
        H.aopen = (hid_t(*)(hid_t, const char *, hid_t))sym("H5Aopen");
        H.acreate = (hid_t(*)(hid_t, const char *, hid_t, hid_t, hid_t, hid_t))sym("H5Acreate2");
        H.aread = (herr_t(*)(hid_t, hid_t, void *))sym("H5Aread");
        H.awrite = (herr_t(*)(hid_t, hid_t, const void *))sym("H5Awrite");
        H.aclose = (herr_t(*)(hid_t))sym("H5Aclose");
        H.t_int32 = (hid_t *)sym("H5T_NATIVE_INT32_g");
        H.t_double = (hid_t *)sym("H5T_NATIVE_DOUBLE_g");
        if (!good || H.open() < 0) {
            H.why = "libhdf5 loaded but its 1.10 C API is incomplete";
            return;
        }
        H.eset_auto(kE_DEFAULT, nullptr, nullptr);  // errors are reported through status codes
        H.ok = true;
    });
    return H;
}

struct Hid {  // closes an HDF5 handle on scope exit
    hid_t id;
    herr_t (*close)(hid_t);
    ~Hid()
    {
        if (id >= 0 && close) close(id);
    }
};

int64_t extent(H5 &H, hid_t dset, std::vector<hsize_t> *dims_out = nullptr)
{
    Hid sp{H.dget_space(dset), H.sclose};
    if (sp.id < 0) return -1;
    int rank = H.sndims(sp.id);
    if (rank < 0) return -1;
    std::vector<hsize_t> d((size_t)std::max(rank, 1), 1), mx((size_t)std::max(rank, 1), 1);
    if (rank > 0 && H.sdims(sp.id, d.data(), mx.data()) < 0) return -1;
    int64_t tot = 1;
    for (int i = 0; i < rank; i++) tot *= (int64_t)d[(size_t)i];
    if (dims_out) *dims_out = d;
    return tot;
}

int write_dataset(H5 &H, hid_t file, const char *name, hid_t type, int rank, const hsize_t *dims,
                  const void *data, const int32_t *attr = nullptr, int nattr = 0)
{
    Hid sp{H.screate_simple(rank, dims, nullptr), H.sclose};
    if (sp.id < 0) return fail(E3X_EIO, std::string("H5Screate_simple failed for ") + name);
    Hid ds{H.dcreate(file, name, type, sp.id, kP_DEFAULT, kP_DEFAULT, kP_DEFAULT), H.dclose};
    if (ds.id < 0) return fail(E3X_EIO, std::string("H5Dcreate2 failed for ") + name);
    if (H.dwrite(ds.id, type, kS_ALL, kS_ALL, kP_DEFAULT, data) < 0)
        return fail(E3X_EIO, std::string("H5Dwrite failed for ") + name);
    if (attr) {
        hsize_t ad = (hsize_t)nattr;
        Hid asp{H.screate_simple(1, &ad, nullptr), H.sclose};
        Hid at{H.acreate(ds.id, "dimensions", *H.t_int32, asp.id, kP_DEFAULT, kP_DEFAULT), H.aclose};
        if (at.id < 0 || H.awrite(at.id, *H.t_int32, attr) < 0)
            return fail(E3X_EIO, std::string("attribute write failed for ") + name);
    }
    return E3X_OK;
}

// ------------------------------------------------------------------ generator helpers
struct Rng {
    std::mt19937_64 e;
    explicit Rng(uint64_t s) : e(s) {}
    double u01() { return (double)(e() >> 11) * 0x1.0p-53; }                 // [0, 1)
    double normal()
    {
        const double a = 1.0 - u01(), b = u01();                              // a in (0, 1]
        return std::sqrt(-2.0 * std::log(a)) * std::cos(6.283185307179586 * b);
    }
};

inline double mind(double d, double n)
{
    d = std::fabs(d);
    return d > 0.5 * n ? n - d : d;
}

// Rotation matrix -> quaternion (w, x, y, z), w >= 0 (Shepperd's method).
void mat_to_quat(const double m[3][3], double q[4])
{
    const double tr = m[0][0] + m[1][1] + m[2][2];
    if (tr > 0) {
        double s = std::sqrt(tr + 1.0) * 2.0;
        q[0] = 0.25 * s;
        q[1] = (m[2][1] - m[1][2]) / s;
        q[2] = (m[0][2] - m[2][0]) / s;
        q[3] = (m[1][0] - m[0][1]) / s;
    } else if (m[0][0] > m[1][1] && m[0][0] > m[2][2]) {
        double s = std::sqrt(1.0 + m[0][0] - m[1][1] - m[2][2]) * 2.0;
        q[0] = (m[2][1] - m[1][2]) / s;
        q[1] = 0.25 * s;
        q[2] = (m[0][1] + m[1][0]) / s;
        q[3] = (m[0][2] + m[2][0]) / s;
    } else if (m[1][1] > m[2][2]) {
        double s = std::sqrt(1.0 + m[1][1] - m[0][0] - m[2][2]) * 2.0;
        q[0] = (m[0][2] - m[2][0]) / s;
        q[1] = (m[0][1] + m[1][0]) / s;
        q[2] = 0.25 * s;
        q[3] = (m[1][2] + m[2][1]) / s;
    } else {
        double s = std::sqrt(1.0 + m[2][2] - m[0][0] - m[1][1]) * 2.0;
        q[0] = (m[1][0] - m[0][1]) / s;
        q[1] = (m[0][2] + m[2][0]) / s;
        q[2] = (m[1][2] + m[2][1]) / s;
        q[3] = 0.25 * s;
    }
    if (q[0] < 0)
        for (int i = 0; i < 4; i++) q[i] = -q[i];
}

// Passive Bunge matrix g = Rz(phi2) Rx(Phi) Rz(phi1).
void bunge_matrix(double p1, double P, double p2, double g[3][3])
{
    const double c1 = std::cos(p1), s1 = std::sin(p1), c = std::cos(P), s = std::sin(P), c2 = std::cos(p2),
                 s2 = std::sin(p2);
    g[0][0] = c1 * c2 - s1 * s2 * c;
    g[0][1] = s1 * c2 + c1 * s2 * c;
    g[0][2] = s2 * s;
    g[1][0] = -c1 * s2 - s1 * c2 * c;
    g[1][1] = -s1 * s2 + c1 * c2 * c;
    g[1][2] = c2 * s;
    g[2][0] = s1 * s;
    g[2][1] = -c1 * s;
    g[2][2] = c;
}

// The reference's orientation matrix Q(r) (f90:2538-2570), for the inverse map.
void rodrigues_matrix(const double r[3], double q[3][3])
{
    const double rr = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
    const double t1 = 1.0 - rr, t2 = 1.0 / (1.0 + rr);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) q[i][j] = ((i == j ? t1 : 0.0) + 2.0 * r[i] * r[j]) * t2;
    q[0][1] -= 2.0 * r[2] * t2; q[0][2] += 2.0 * r[1] * t2; q[1][2] -= 2.0 * r[0] * t2;
    q[1][0] += 2.0 * r[2] * t2; q[2][0] -= 2.0 * r[1] * t2; q[2][1] += 2.0 * r[0] * t2;
}

void matrix_to_bunge(const double g[3][3], double e[3])
{
    const double P = std::acos(std::max(-1.0, std::min(1.0, g[2][2])));
    double p1, p2;
    if (std::fabs(std::sin(P)) > 1e-12) {
        p1 = std::atan2(g[2][0], -g[2][1]);
        p2 = std::atan2(g[0][2], g[1][2]);
    } else {
        p1 = std::atan2(g[0][1], g[0][0]);  // phi1 + phi2 (or difference); put it all in phi1
        p2 = 0.0;
    }
    const double tw = 6.283185307179586;
    if (p1 < 0) p1 += tw;
    if (p2 < 0) p2 += tw;
    e[0] = p1; e[1] = P; e[2] = p2;
}

}  // namespace

// ------------------------------------------------------------------ JSON controls
// The controls of a serving request: one flat JSON object whose values are numbers,
// strings, booleans, null or arrays of numbers.
struct JVal {
    enum Kind { kNum, kStr, kBool, kNull, kArr } kind = kNull;
    double num = 0;
    std::string str;
    std::vector<double> arr;
};

struct JParser {
    const char *p;
    std::string err;
    void ws()
    {
        while (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r') p++;
    }
    bool str(std::string &out)
    {
        if (*p != '"') return err = "expected a string", false;
        p++;
        out.clear();
        while (*p && *p != '"') {
            if (*p == '\\') {
                p++;
                if (!*p) break;
                const char c = *p;
                out += c == 'n' ? '\n' : c == 't' ? '\t' : c;  // \" \\ \/ and the rest as written
                p++;
                continue;
            }
            out += *p++;
        }
        if (*p != '"') return err = "unterminated string", false;
        p++;
        return true;
    }
    bool num(double &out)
    {
        char *end = nullptr;
        out = std::strtod(p, &end);
        if (end == p) return err = "expected a number", false;
        p = end;
        return true;
    }
    bool value(JVal &v)
    {
        ws();
        if (*p == '"') {
            v.kind = JVal::kStr;
            return str(v.str);
        }
        if (*p == '[') {
            p++;
            v.kind = JVal::kArr;
            ws();
            if (*p == ']') {
                p++;
                return true;
            }
            while (true) {
                ws();
                double d;
                if (!num(d)) return false;
                v.arr.push_back(d);
                ws();
                if (*p == ',') {
                    p++;
                    continue;
                }
                if (*p == ']') {
                    p++;
                    return true;
                }
                return err = "expected , or ] in an array", false;
            }
        }
        if (!std::strncmp(p, "true", 4)) { p += 4; v.kind = JVal::kBool; v.num = 1; return true; }
        if (!std::strncmp(p, "false", 5)) { p += 5; v.kind = JVal::kBool; v.num = 0; return true; }
        if (!std::strncmp(p, "null", 4)) { p += 4; v.kind = JVal::kNull; return true; }
        v.kind = JVal::kNum;
        return num(v.num);
    }
    bool object(std::map<std::string, JVal> &out)
    {
        ws();
        if (*p != '{') return err = "expected a JSON object", false;
        p++;
        ws();
        if (*p == '}') {
            p++;
            return true;
        }
        while (true) {
            ws();
            std::string key;
            if (!str(key)) return false;
            ws();
            if (*p != ':') return err = "expected : after a key", false;
            p++;
            JVal v;
            if (!value(v)) return false;
            out[key] = v;
            ws();
            if (*p == ',') {
                p++;
                continue;
            }
            if (*p == '}') {
                p++;
                ws();
                if (*p) return err = "text after the object", false;
                return true;
            }
            return err = "expected , or } in the object", false;
        }
    }
};

extern "C" {

const char *e3x_io_last_error(void) { return g_err.c_str(); }

int e3x_h5_probe_input(const char *path, int64_t *npix, int64_t *norient, int32_t dims[3])
{
    H5 &H = h5();
    if (!H.ok) return fail(E3X_EIO, H.why);
    if (!path) return fail(E3X_EINVAL, "path is NULL");
    Hid f{H.fopen(path, kF_ACC_RDONLY, kP_DEFAULT), H.fclose};
    if (f.id < 0) return fail(E3X_EIO, std::string("Error opening HDF5 file ") + path);
    Hid dp{H.dopen(f.id, "pix", kP_DEFAULT), H.dclose};
    if (dp.id < 0) return fail(E3X_EIO, "dataset /pix missing");
    if (npix) *npix = extent(H, dp.id);
    if (dims) {
        dims[0] = dims[1] = dims[2] = 0;
        if (H.aexists(dp.id, "dimensions") > 0) {
            Hid at{H.aopen(dp.id, "dimensions", kP_DEFAULT), H.aclose};
            int32_t d[3] = {0, 0, 0};
            if (at.id >= 0 && H.aread(at.id, *H.t_int32, d) >= 0) std::memcpy(dims, d, sizeof d);
        }
    }
    Hid dor{H.dopen(f.id, "orientation", kP_DEFAULT), H.dclose};
    if (norient) *norient = dor.id < 0 ? 0 : extent(H, dor.id);
    return E3X_OK;
}

int e3x_h5_read_input(const char *path, int64_t ns, int nphase, int32_t *pix, double *orientation)
{
    H5 &H = h5();
    if (!H.ok) return fail(E3X_EIO, H.why);
    if (!path || !pix) return fail(E3X_EINVAL, "path or pix is NULL");
    Hid f{H.fopen(path, kF_ACC_RDONLY, kP_DEFAULT), H.fclose};
    if (f.id < 0) return fail(E3X_EIO, std::string("Error opening HDF5 file ") + path);
    {
        Hid d{H.dopen(f.id, "pix", kP_DEFAULT), H.dclose};
        if (d.id < 0) return fail(E3X_EIO, "dataset /pix missing");
        const int64_t n = extent(H, d.id);
        if (n != ns)
            return fail(E3X_EINVAL, "HDF5 pix size (" + std::to_string(n) + ") does not match simulation ns (" +
                                        std::to_string(ns) + ")");
        if (H.dread(d.id, *H.t_int32, kS_ALL, kS_ALL, kP_DEFAULT, pix) < 0) return fail(E3X_EIO, "reading /pix failed");
    }
    if (nphase > 1) {
        Hid d{H.dopen(f.id, "orientation", kP_DEFAULT), H.dclose};
        if (d.id < 0) return fail(E3X_EIO, "dataset /orientation missing");
        const int64_t n = extent(H, d.id);
        if (n != 3 * (int64_t)(nphase - 1))
            return fail(E3X_EINVAL, "HDF5 orientation size (" + std::to_string(n) + ") mismatch with allocated (" +
                                        std::to_string(3 * (int64_t)(nphase - 1)) + ")");
        if (orientation && H.dread(d.id, *H.t_double, kS_ALL, kS_ALL, kP_DEFAULT, orientation) < 0)
            return fail(E3X_EIO, "reading /orientation failed");
    }
    int32_t lo = INT32_MAX, hi = INT32_MIN;
    for (int64_t m = 0; m < ns; m++) {
        lo = std::min(lo, pix[m]);
        hi = std::max(hi, pix[m]);
    }
    if (lo < 1 || hi > nphase) return fail(E3X_EPHASE, "Phase label in pix out of bounds");
    return E3X_OK;
}

int e3x_h5_write_input(const char *path, int nx, int ny, int nz, const int32_t *pix, int ngrain,
                       const double *orientation)
{
    H5 &H = h5();
    if (!H.ok) return fail(E3X_EIO, H.why);
    if (!path || !pix || (ngrain > 0 && !orientation)) return fail(E3X_EINVAL, "NULL argument");
    Hid f{H.fcreate(path, kF_ACC_TRUNC, kP_DEFAULT, kP_DEFAULT), H.fclose};
    if (f.id < 0) return fail(E3X_EIO, std::string("cannot create ") + path);
    const hsize_t ns = (hsize_t)nx * ny * nz;
    const hsize_t dp[2] = {1, ns};
    const int32_t ad[3] = {nx, ny, nz};
    if (int rc = write_dataset(H, f.id, "pix", *H.t_int32, 2, dp, pix, ad, 3)) return rc;
    const hsize_t dor[2] = {1, (hsize_t)3 * ngrain};
    const int32_t ao[2] = {3 * ngrain, 1};
    if (int rc = write_dataset(H, f.id, "orientation", *H.t_double, 2, dor, orientation, ao, 2)) return rc;
    return E3X_OK;
}

int e3x_h5_write_fullfield(const char *path, int64_t ns, const int32_t *pix, const double *vm,
                           const double *stress)
{
    H5 &H = h5();
    if (!H.ok) return fail(E3X_EIO, H.why);
    if (!path || !pix || !vm || !stress) return fail(E3X_EINVAL, "NULL argument");
    Hid f{H.fcreate(path, kF_ACC_TRUNC, kP_DEFAULT, kP_DEFAULT), H.fclose};
    if (f.id < 0) return fail(E3X_EIO, std::string("cannot create ") + path);
    const hsize_t d1 = (hsize_t)ns;
    if (int rc = write_dataset(H, f.id, "pix", *H.t_int32, 1, &d1, pix)) return rc;
    if (int rc = write_dataset(H, f.id, "vm", *H.t_double, 1, &d1, vm)) return rc;
    const hsize_t d2[2] = {6, (hsize_t)ns};  // Fortran dims (ns, 6) reversed
    if (int rc = write_dataset(H, f.id, "stress_tensor", *H.t_double, 2, d2, stress)) return rc;
    return E3X_OK;
}

int e3x_h5_read_fullfield(const char *path, int64_t ns, int32_t *pix, double *vm, double *stress)
{
    H5 &H = h5();
    if (!H.ok) return fail(E3X_EIO, H.why);
    Hid f{H.fopen(path, kF_ACC_RDONLY, kP_DEFAULT), H.fclose};
    if (f.id < 0) return fail(E3X_EIO, std::string("cannot open ") + path);
    struct Item { const char *name; hid_t type; void *dst; int64_t n; };
    const Item items[3] = {{"pix", *H.t_int32, pix, ns}, {"vm", *H.t_double, vm, ns},
                           {"stress_tensor", *H.t_double, stress, 6 * ns}};
    for (const Item &it : items) {
        if (!it.dst) continue;
        Hid d{H.dopen(f.id, it.name, kP_DEFAULT), H.dclose};
        if (d.id < 0) return fail(E3X_EIO, std::string("dataset missing: ") + it.name);
        if (extent(H, d.id) != it.n) return fail(E3X_EINVAL, std::string("size mismatch: ") + it.name);
        if (H.dread(d.id, it.type, kS_ALL, kS_ALL, kP_DEFAULT, it.dst) < 0)
            return fail(E3X_EIO, std::string("read failed: ") + it.name);
    }
    return E3X_OK;
}

int e3x_rodrigues_from_bunge(int n, const double *euler, double *rod)
{
    if (n < 0 || (n > 0 && (!euler || !rod))) return fail(E3X_EINVAL, "bad arguments");
    for (int g = 0; g < n; g++) {
        double m[3][3], q[4];
        bunge_matrix(euler[3 * g], euler[3 * g + 1], euler[3 * g + 2], m);
        mat_to_quat(m, q);
        const double w = std::fabs(q[0]) < 1e-8 ? 1e-9 : q[0];  // as microstructure_gen.m:793-797
        for (int c = 0; c < 3; c++) rod[3 * g + c] = q[1 + c] / w;
    }
    return E3X_OK;
}

int e3x_bunge_from_rodrigues(int n, const double *rod, double *euler)
{
    if (n < 0 || (n > 0 && (!euler || !rod))) return fail(E3X_EINVAL, "bad arguments");
    for (int g = 0; g < n; g++) {
        double q[3][3];
        rodrigues_matrix(rod + 3 * g, q);
        matrix_to_bunge(q, euler + 3 * g);
    }
    return E3X_OK;
}

int e3x_generate(const e3x_gen_params *p, int32_t *pix, double *rod, double *euler, double *pore_out)
{
    if (!p || !pix || !rod) return fail(E3X_EINVAL, "NULL argument");
    const int nx = p->nx, ny = p->ny, nz = p->nz, ng = p->ngrain;
    if (nx < 1 || ny < 1 || nz < 1 || ng < 1) return fail(E3X_EINVAL, "grid and ngrain must be positive");
    if (p->pore_fraction < 0 || p->pore_fraction >= 1 || p->lof_fraction < 0 || p->lof_fraction >= 1)
        return fail(E3X_EINVAL, "pore fractions must be in [0, 1)");
    const int64_t ns = (int64_t)nx * ny * nz;
    Rng rng(p->seed);
    // grain centres, then orientations (fixed draw order)
    std::vector<double> sx((size_t)ng), sy((size_t)ng), sz((size_t)ng);
    for (int g = 0; g < ng; g++) {
        sx[(size_t)g] = rng.u01() * nx;
        sy[(size_t)g] = rng.u01() * ny;
        sz[(size_t)g] = rng.u01() * nz;
    }
    for (int g = 0; g < ng; g++) {
        const double u1 = rng.u01(), u2 = rng.u01(), u3 = rng.u01();
        const double tw = 6.283185307179586;
        const double qw = std::sqrt(1 - u1) * std::sin(tw * u2), qx = std::sqrt(1 - u1) * std::cos(tw * u2);
        const double qy = std::sqrt(u1) * std::sin(tw * u3), qz = std::sqrt(u1) * std::cos(tw * u3);
        const double w = std::fabs(qw) < 1e-8 ? 1e-9 : qw;
        rod[3 * g] = qx / w;
        rod[3 * g + 1] = qy / w;
        rod[3 * g + 2] = qz / w;
    }
    if (euler) e3x_bunge_from_rodrigues(ng, rod, euler);

    // periodic Voronoi via a cell list searched in Chebyshev rings
    const double nbar = std::cbrt((double)ns);
    const double per = std::cbrt(std::max(1.0, ng / 2.0));
    const int G[3] = {std::max(1, std::min(nx, (int)(per * nx / nbar))), std::max(1, std::min(ny, (int)(per * ny / nbar))),
                      std::max(1, std::min(nz, (int)(per * nz / nbar)))};
    const double cs[3] = {(double)nx / G[0], (double)ny / G[1], (double)nz / G[2]};
    const double csmin = std::min(cs[0], std::min(cs[1], cs[2]));
    const int ncell = G[0] * G[1] * G[2];
    std::vector<int> cstart((size_t)ncell + 1, 0), citem((size_t)ng);
    auto cell_of = [&](int g) {
        int cx = std::min(G[0] - 1, (int)(sx[(size_t)g] / cs[0]));
        int cy = std::min(G[1] - 1, (int)(sy[(size_t)g] / cs[1]));
        int cz = std::min(G[2] - 1, (int)(sz[(size_t)g] / cs[2]));
        return cx + G[0] * (cy + G[1] * cz);
    };
    for (int g = 0; g < ng; g++) cstart[(size_t)cell_of(g) + 1]++;
    for (int c = 0; c < ncell; c++) cstart[(size_t)c + 1] += cstart[(size_t)c];
    {
        std::vector<int> fill(cstart.begin(), cstart.end() - 1);
        for (int g = 0; g < ng; g++) citem[(size_t)fill[(size_t)cell_of(g)]++] = g;
    }
    const int gmax = std::max(G[0], std::max(G[1], G[2]));
#pragma omp parallel
    {
        std::vector<int64_t> stamp((size_t)ncell, -1);
#pragma omp for schedule(static)
        for (int64_t m = 0; m < ns; m++) {
            const double px = (double)(m % nx) + 0.5, py = (double)((m / nx) % ny) + 0.5,
                         pz = (double)(m / ((int64_t)nx * ny)) + 0.5;
            const int hx = std::min(G[0] - 1, (int)(px / cs[0])), hy = std::min(G[1] - 1, (int)(py / cs[1])),
                      hz = std::min(G[2] - 1, (int)(pz / cs[2]));
            double best = INFINITY;
            int bg = -1;
            for (int r = 0;; r++) {
                for (int dz = -r; dz <= r; dz++)
                    for (int dy = -r; dy <= r; dy++)
                        for (int dx = -r; dx <= r; dx++) {
                            if (std::max(std::abs(dx), std::max(std::abs(dy), std::abs(dz))) != r) continue;
                            const int cx = ((hx + dx) % G[0] + G[0]) % G[0], cy = ((hy + dy) % G[1] + G[1]) % G[1],
                                      cz = ((hz + dz) % G[2] + G[2]) % G[2];
                            const int c = cx + G[0] * (cy + G[1] * cz);
                            if (stamp[(size_t)c] == m) continue;
                            stamp[(size_t)c] = m;
                            for (int t = cstart[(size_t)c]; t < cstart[(size_t)c + 1]; t++) {
                                const int g = citem[(size_t)t];
                                const double ddx = mind(px - sx[(size_t)g], nx), ddy = mind(py - sy[(size_t)g], ny),
                                             ddz = mind(pz - sz[(size_t)g], nz);
                                const double d2 = ddx * ddx + ddy * ddy + ddz * ddz;
                                if (d2 < best || (d2 == best && g < bg)) {
                                    best = d2;
                                    bg = g;
                                }
                            }
                        }
                const double reach = r * csmin;
                if ((bg >= 0 && best < reach * reach) || 2 * r + 1 >= gmax) break;
            }
            pix[m] = bg + 1;
        }
    }

    // pores
    const int32_t pore = ng + 1;
    int64_t npore = 0;
    auto mark = [&](int64_t m) {
        if (pix[m] != pore) {
            pix[m] = pore;
            npore++;
        }
    };
    auto paint = [&](double cx, double cy, double cz, double ra, double rb, double rc, double yaw) {
        const double cyaw = std::cos(yaw), syaw = std::sin(yaw);
        const double ext = std::max(ra, std::max(rb, rc));
        const int x0 = (int)std::floor(cx - ext - 1), x1 = (int)std::ceil(cx + ext + 1);
        const int y0 = (int)std::floor(cy - ext - 1), y1 = (int)std::ceil(cy + ext + 1);
        const int z0 = (int)std::floor(cz - rc - 1), z1 = (int)std::ceil(cz + rc + 1);
        for (int k = z0; k <= z1; k++)
            for (int j = y0; j <= y1; j++)
                for (int i = x0; i <= x1; i++) {
                    const double dx = i + 0.5 - cx, dy = j + 0.5 - cy, dz = k + 0.5 - cz;
                    const double xr = cyaw * dx + syaw * dy, yr = -syaw * dx + cyaw * dy;
                    if ((xr * xr) / (ra * ra) + (yr * yr) / (rb * rb) + (dz * dz) / (rc * rc) > 1.0) continue;
                    const int ii = ((i % nx) + nx) % nx, jj = ((j % ny) + ny) % ny, kk = ((k % nz) + nz) % nz;
                    mark(ii + (int64_t)nx * (jj + (int64_t)ny * kk));
                }
    };
    const int64_t target = (int64_t)std::llround(p->pore_fraction * (double)ns);
    if (p->pore_mode == E3X_PORES_VOXELS) {
        while (npore < target) mark((int64_t)(rng.u01() * (double)ns) % ns);
    } else if (p->pore_mode == E3X_PORES_SPHERES) {
        if (!(p->pore_median > 0)) return fail(E3X_EINVAL, "pore_median must be positive");
        while (npore < target) {
            const double cx = rng.u01() * nx, cy = rng.u01() * ny, cz = rng.u01() * nz;
            const double r = p->pore_median * std::exp(p->pore_sigma * rng.normal());
            paint(cx, cy, cz, r, r, r, 0.0);
        }
    } else if (p->pore_mode != E3X_PORES_NONE) {
        return fail(E3X_EINVAL, "unknown pore_mode");
    }
    if (p->lof_fraction > 0) {
        if (!(p->lof_axes[0] > 0 && p->lof_axes[1] > 0 && p->lof_axes[2] > 0))
            return fail(E3X_EINVAL, "lof_axes must be positive");
        const int64_t lof_target = npore + (int64_t)std::llround(p->lof_fraction * (double)ns);
        while (npore < lof_target) {
            const double cx = rng.u01() * nx, cy = rng.u01() * ny, cz = rng.u01() * nz;
            const double yaw = rng.u01() * 6.283185307179586;
            paint(cx, cy, cz, p->lof_axes[0], p->lof_axes[1], p->lof_axes[2], yaw);
        }
    }
    if (pore_out) *pore_out = (double)npore / (double)ns;
    return E3X_OK;
}

int e3x_request_from_files(const char *h5_path, const char *json, e3x_request *rq, int32_t **pix,
                           double **orientation, char *fullfield_path, int path_cap)
{
    if (!h5_path || !rq || !pix || !orientation) return fail(E3X_EINVAL, "NULL argument");
    *pix = nullptr;
    *orientation = nullptr;
    if (fullfield_path && path_cap > 0) fullfield_path[0] = 0;
    std::map<std::string, JVal> c;
    if (json && *json) {
        JParser jp{json, {}};
        if (!jp.object(c)) return fail(E3X_EINVAL, "JSON controls: " + jp.err);
    }
    static const char *known[] = {"grid", "c11", "c12", "c44", "zener", "strain6", "phasemod", "flag_m",
                                  "precision", "replace_every", "tol", "ldemb", "fullfield"};
    for (const auto &kv : c) {
        bool ok = false;
        for (const char *k : known) ok |= kv.first == k;
        if (!ok) return fail(E3X_EINVAL, "JSON controls: unknown key \"" + kv.first + "\"");
    }
    auto num = [&](const char *k, double dflt, double *out) -> int {
        auto it = c.find(k);
        if (it == c.end() || it->second.kind == JVal::kNull) { *out = dflt; return E3X_OK; }
        if (it->second.kind != JVal::kNum && it->second.kind != JVal::kBool)
            return fail(E3X_EINVAL, std::string("JSON controls: ") + k + " must be a number");
        *out = it->second.num;
        return E3X_OK;
    };
    auto arr = [&](const char *k, size_t n, std::vector<double> *out) -> int {
        auto it = c.find(k);
        if (it == c.end()) return E3X_OK;
        if (it->second.kind != JVal::kArr || it->second.arr.size() != n)
            return fail(E3X_EINVAL, std::string("JSON controls: ") + k + " must be an array of " + std::to_string(n) +
                                        " numbers");
        *out = it->second.arr;
        return E3X_OK;
    };
    // the grid: "grid", else the /pix dimensions attribute the MATLAB generator writes
    int64_t npix = 0, norient = 0;
    int32_t dims[3] = {0, 0, 0};
    if (int rc = e3x_h5_probe_input(h5_path, &npix, &norient, dims)) return rc;
    std::vector<double> grid;
    if (int rc = arr("grid", 3, &grid)) return rc;
    int nx = dims[0], ny = dims[1], nz = dims[2];
    if (!grid.empty()) {
        nx = (int)grid[0];
        ny = (int)grid[1];
        nz = (int)grid[2];
    }
    if (nx < 2 || ny < 2 || nz < 2)
        return fail(E3X_EINVAL, "grid unknown: give \"grid\" (the file has no /pix dimensions attribute)");
    const int64_t ns = (int64_t)nx * ny * nz;
    if (norient % 3 != 0) return fail(E3X_EINVAL, "/orientation holds a number of values not divisible by 3");
    const int nphase = (int)(norient / 3) + 1;  // n_grains + 1: the pore phase (f90:88)
    // materials: Pa; "zener" A sets c11 = c12 + 2 c44 / A
    double c11, c12, c44, zener;
    if (int rc = num("c12", 147.3e9, &c12)) return rc;
    if (int rc = num("c44", 124.7e9, &c44)) return rc;
    if (int rc = num("zener", 0.0, &zener)) return rc;
    if (int rc = num("c11", zener > 0 ? c12 + 2.0 * c44 / zener : 246.5e9, &c11)) return rc;
    if (c.count("c11") && zener > 0) return fail(E3X_EINVAL, "JSON controls: give c11 or zener, not both");
    std::vector<double> e6 = {(double)1.0e-3f, 0, 0, 0, 0, 0}, pm = {0, 0};  // aml = 1.0e-3 (f90:125)
    if (int rc = arr("strain6", 6, &e6)) return rc;
    if (c.count("phasemod")) {
        auto it = c.find("phasemod");
        if (it->second.kind != JVal::kArr || (it->second.arr.size() != 2 && it->second.arr.size() != 4))
            return fail(E3X_EINVAL, "JSON controls: phasemod must hold 2 or 4 numbers");
        pm = it->second.arr;
    }
    double flag_m, re, tol, ldemb;
    if (int rc = num("flag_m", 1, &flag_m)) return rc;
    if (int rc = num("replace_every", 50, &re)) return rc;
    if (int rc = num("tol", (double)1.0e-8f, &tol)) return rc;  // tol = 1.0d-8 as real (f90:148)
    if (int rc = num("ldemb", 10000, &ldemb)) return rc;        // f90:141
    int precision = E3X_PREC_FP64;
    if (auto it = c.find("precision"); it != c.end()) {
        if (it->second.kind != JVal::kStr || (it->second.str != "fp64" && it->second.str != "mixed"))
            return fail(E3X_EINVAL, "JSON controls: precision must be \"fp64\" or \"mixed\"");
        precision = it->second.str == "mixed" ? E3X_PREC_MIXED : E3X_PREC_FP64;
    }
    if (auto it = c.find("fullfield"); it != c.end() && it->second.kind != JVal::kNull) {
        if (it->second.kind != JVal::kStr) return fail(E3X_EINVAL, "JSON controls: fullfield must be a path");
        if (!fullfield_path || (int)it->second.str.size() + 1 > path_cap)
            return fail(E3X_EINVAL, "fullfield path does not fit the caller's buffer");
        std::memcpy(fullfield_path, it->second.str.c_str(), it->second.str.size() + 1);
    }
    // the microstructure, with the reference's checks (f90:2879-2991)
    int32_t *p = static_cast<int32_t *>(std::malloc(sizeof(int32_t) * (size_t)ns));
    double *o = static_cast<double *>(std::malloc(sizeof(double) * (size_t)std::max<int64_t>(norient, 1)));
    if (!p || !o) {
        std::free(p);
        std::free(o);
        return fail(E3X_ENOMEM, "host allocation failed");
    }
    if (int rc = e3x_h5_read_input(h5_path, ns, nphase, p, o)) {
        std::free(p);
        std::free(o);
        return rc;
    }
    std::memset(rq, 0, sizeof(*rq));
    rq->nx = nx; rq->ny = ny; rq->nz = nz; rq->nphase = nphase;
    rq->pix = p;
    rq->orientation = nphase > 1 ? o : nullptr;
    rq->c11 = c11; rq->c12 = c12; rq->c44 = c44;
    for (size_t k = 0; k < pm.size(); k++) rq->phasemod[k] = pm[k];
    rq->flag_m = (int)flag_m;
    for (int k = 0; k < 6; k++) rq->strain6[k] = e6[(size_t)k];
    rq->precision = precision;
    rq->replace_every = precision == E3X_PREC_MIXED ? (int)re : 0;
    rq->tol = (float)tol;
    rq->ldemb = (int)ldemb;
    *pix = p;
    *orientation = o;
    return E3X_OK;
}

void e3x_request_free(int32_t *pix, double *orientation)
{
    std::free(pix);
    std::free(orientation);
}

}  // extern "C"
