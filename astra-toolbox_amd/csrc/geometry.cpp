// Host-side geometry precomputation (FP64) for the device kernels, plus the
// small runtime helpers (errors, scratch allocation).
//
// The per-angle coefficients follow the reference's transferConstants
// routines; they are computed once per geometry handle (not once per call as
// in the reference) and kept in device memory.
#include <cmath>
#include <complex>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>

#include "internal.hpp"

namespace amx {

void fail(int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    throw Error(code, buf);
}

// ------------------------------------------------------------- memory ----
void configure_pool_once() {
    static std::mutex m;
    static std::vector<int> done;
    int dev = 0;
    AMX_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(m);
    for (int d : done)
        if (d == dev) return;
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
        uint64_t thr = UINT64_MAX;
        (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
    }
    (void)hipGetLastError();
    done.push_back(dev);
}

void scratch_usage(size_t &in_use, size_t &peak, bool reset_peak) {
    int dev = 0;
    AMX_HIP(hipGetDevice(&dev));
    hipMemPool_t pool;
    AMX_HIP(hipDeviceGetDefaultMemPool(&pool, dev));
    uint64_t cur = 0, high = 0;
    AMX_HIP(hipMemPoolGetAttribute(pool, hipMemPoolAttrUsedMemCurrent, &cur));
    AMX_HIP(hipMemPoolGetAttribute(pool, hipMemPoolAttrUsedMemHigh, &high));
    in_use = (size_t)cur;
    peak = (size_t)high;
    if (reset_peak) {
        uint64_t zero = 0;
        AMX_HIP(hipMemPoolSetAttribute(pool, hipMemPoolAttrUsedMemHigh, &zero));
    }
}

Scratch::Scratch(size_t bytes, hipStream_t s) : s_(s) {
    configure_pool_once();
    if (bytes == 0) bytes = 256;
    hipError_t e = hipMallocAsync(&p_, bytes, s);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        fail(ASTRA_AMD_ERR_ALLOC, "device scratch allocation of %zu bytes failed: %s", bytes,
             hipGetErrorString(e));
    }
}

Scratch::~Scratch() {
    if (p_) (void)hipFreeAsync(p_, s_);
}

DevArray::DevArray(size_t bytes) : bytes_(bytes) {
    if (bytes == 0) return;
    hipError_t e = hipMalloc(&p_, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        p_ = nullptr;
        fail(ASTRA_AMD_ERR_ALLOC, "device allocation of %zu bytes failed: %s", bytes,
             hipGetErrorString(e));
    }
}

DevArray::~DevArray() {
    if (p_) (void)hipFree(p_);
}

DevArray &DevArray::operator=(DevArray &&o) noexcept {
    if (this != &o) {
        if (p_) (void)hipFree(p_);
        p_ = o.p_;
        bytes_ = o.bytes_;
        o.p_ = nullptr;
        o.bytes_ = 0;
    }
    return *this;
}

void DevArray::upload(const void *host, size_t bytes) {
    AMX_REQUIRE(bytes <= bytes_, "internal: upload larger than buffer");
    AMX_HIP(hipMemcpy(p_, host, bytes, hipMemcpyHostToDevice));
}

unsigned next_pow2(unsigned n) {
    // calcNextPowerOfTwo (src/Filters.cpp:662-669)
    unsigned x = 1;
    while (x < n) x *= 2;
    return x;
}

// ----------------------------------------------------------- 3D ----------
namespace {

struct V3 {
    double x, y, z;
};
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline double det3x(V3 b, V3 c) { return b.y * c.z - b.z * c.y; }
inline double det3y(V3 b, V3 c) { return -(b.x * c.z - b.z * c.x); }
inline double det3z(V3 b, V3 c) { return b.x * c.y - b.y * c.x; }
inline double det3(V3 a, V3 b, V3 c) {
    return a.x * det3x(b, c) + a.y * det3y(b, c) + a.z * det3z(b, c);
}
// norm of cross(a, b) with components scaled by (sy sz, sx sz, sx sy)
inline double scaled_cross_norm(V3 a, V3 b, const float *sc) {
    double cx = det3x(a, b) * sc[1] * sc[2];
    double cy = det3y(a, b) * sc[0] * sc[2];
    double cz = det3z(a, b) * sc[0] * sc[1];
    return std::sqrt(cx * cx + cy * cy + cz * cz);
}

const int kPerm[3][3] = {{0, 1, 2}, {1, 0, 2}, {2, 0, 1}};

int principal_axis(float dX, float dY, float dZ) {
    // ties resolve to the lower axis (cuda/3d/cone_fp.cu:383-388)
    if (dX >= dY && dX >= dZ) return 0;
    if (dY >= dX && dY >= dZ) return 1;
    return 2;
}

void cone_bp_coeffs(const double *p, int kernel, const float *vs, BP3DAngle &o) {
    // DevConeParams of cuda/3d/cone_bp.cu:221-302, with the texel-centre
    // offset (-0.5) and the 2-texel padding (+2) folded into the numerators.
    V3 s{p[0], p[1], p[2]}, d{p[3], p[4], p[5]}, u{p[6], p[7], p[8]}, v{p[9], p[10], p[11]};
    V3 smd = sub(s, d);
    double sc;
    if (kernel == ASTRA_AMD_KER3D_FDK_WEIGHTING)
        sc = 1.0 / det3(u, v, s);
    else if (kernel == ASTRA_AMD_KER3D_2D_WEIGHTING)
        sc = scaled_cross_norm(u, V3{0, 0, 1}, vs) / det3(u, v, smd);
    else
        sc = std::sqrt(scaled_cross_norm(u, v, vs)) / det3(u, v, smd);
    double nu[4] = {sc * det3x(v, smd), sc * det3y(v, smd), sc * det3z(v, smd), sc * det3(s, v, d)};
    double nv[4] = {-sc * det3x(u, smd), -sc * det3y(u, smd), -sc * det3z(u, smd),
                    -sc * det3(s, u, d)};
    double dn[4] = {-sc * det3x(u, v), -sc * det3y(u, v), -sc * det3z(u, v), sc * det3(u, v, s)};
    if (kernel == ASTRA_AMD_KER3D_FDK_WEIGHTING) dn[3] = 1.0;  // cone_bp.cu:101
    for (int k = 0; k < 4; ++k) {
        o.nu[k] = (float)(nu[k] + 1.5 * dn[k]);
        o.nv[k] = (float)(nv[k] + 1.5 * dn[k]);
        o.dn[k] = (float)dn[k];
    }
}

void par3d_bp_coeffs(const double *p, int kernel, const float *vs, BP3DAngle &o, float &w) {
    // DevPar3DParams + scale of cuda/3d/par3d_bp.cu:202-244
    V3 r{p[0], p[1], p[2]}, d{p[3], p[4], p[5]}, u{p[6], p[7], p[8]}, v{p[9], p[10], p[11]};
    double den = det3(r, u, v);
    double nu[4] = {-det3x(r, v) / den, -det3y(r, v) / den, -det3z(r, v) / den,
                    -det3(r, d, v) / den};
    double nv[4] = {det3x(r, u) / den, det3y(r, u) / den, det3z(r, u) / den, det3(r, d, u) / den};
    for (int k = 0; k < 4; ++k) {
        o.nu[k] = (float)(nu[k] + (k == 3 ? 1.5 : 0.0));
        o.nv[k] = (float)(nv[k] + (k == 3 ? 1.5 : 0.0));
        o.dn[k] = (k == 3) ? 1.0f : 0.0f;
    }
    if (kernel == ASTRA_AMD_KER3D_2D_WEIGHTING)
        w = (float)(1.0 / scaled_cross_norm(u, V3{0, 0, 1}, vs));
    else
        w = (float)(1.0 / scaled_cross_norm(u, v, vs));
}

inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double norm(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 scale(V3 a, double f) { return {a.x * f, a.y * f, a.z * f}; }
inline V3 cross(V3 a, V3 b) { return {det3x(a, b), det3y(a, b), det3z(a, b)}; }

// Cylinder axes of one cyl_cone_vec row [src, detC, u, v, R]
// (getCylConeAxes, src/GeometryUtil3D.cpp:542-561): radial a from the axis
// to the detector centre (away from the source), tangential b, centre c.
void cyl_axes(const double *p, V3 &a, V3 &b, V3 &c) {
    V3 s{p[0], p[1], p[2]}, d{p[3], p[4], p[5]}, u{p[6], p[7], p[8]}, v{p[9], p[10], p[11]};
    const double R = p[12];
    a = scale(cross(u, v), -R / (norm(u) * norm(v)));
    V3 dm{d.x - a.x - s.x, d.y - a.y - s.y, d.z - a.z - s.z};
    V3 dp{d.x + a.x - s.x, d.y + a.y - s.y, d.z + a.z - s.z};
    if (norm(dm) > norm(dp)) a = scale(a, -1.0);
    c = sub(d, a);
    b = scale(u, R / norm(u));
}

// DevCylConeParamsBP (cuda/3d/cone_cyl_bp.cu:165-230)
void cyl_bp_coeffs(const double *p, CylBPAngle &o) {
    V3 s{p[0], p[1], p[2]}, u{p[6], p[7], p[8]}, v{p[9], p[10], p[11]};
    const double R = p[12];
    V3 a, b, c;
    cyl_axes(p, a, b, c);
    const V3 e[3] = {scale(a, 1.0 / norm(a)), scale(b, 1.0 / norm(b)), scale(v, 1.0 / norm(v))};
    // ray density 1 / t^2 after scaling the ray components by sqrt(|u x v|)
    const double f = std::sqrt(norm(cross(u, v)));
    float *rows[3] = {o.ca, o.cb, o.cc};
    const V3 smc = sub(s, c);
    double sm[3];
    for (int k = 0; k < 3; ++k) {
        rows[k][0] = (float)(e[k].x * f);
        rows[k][1] = (float)(e[k].y * f);
        rows[k][2] = (float)(e[k].z * f);
        rows[k][3] = (float)(-dot(s, e[k]) * f);
        sm[k] = dot(smc, e[k]);
    }
    o.sa = (float)sm[0];
    o.sb = (float)sm[1];
    o.sc = (float)sm[2];
    o.c = (float)(sm[0] * sm[0] + sm[1] * sm[1] - R * R);
    o.inv_dth = (float)(R / norm(u));
    o.inv_nv = (float)(1.0 / norm(v));
    o.inv_r = (float)(1.0 / R);
    o.pad_ = 0.f;
}

}  // namespace

Geom3D *geom3d_build(int kind, unsigned nx, unsigned ny, unsigned nz, unsigned na, unsigned nu,
                     unsigned nv, const double *vecs, const float *vol_scale) {
    AMX_REQUIRE(kind == ASTRA_AMD_GEOM_CONE || kind == ASTRA_AMD_GEOM_PAR3D ||
                    kind == ASTRA_AMD_GEOM_CYLCONE,
                "geom3d: kind must be CONE, PAR3D or CYLCONE (got %d)", kind);
    const bool cyl = kind == ASTRA_AMD_GEOM_CYLCONE;
    const size_t per = cyl ? 13 : 12;
    AMX_REQUIRE(nx > 0 && ny > 0 && nz > 0 && na > 0 && nu > 0 && nv > 0,
                "geom3d: all dimensions must be positive");
    AMX_REQUIRE(vecs != nullptr, "geom3d: vecs is NULL");
    // padded planes must stay below 2^24 texels (offsets are formed in fp32)
    AMX_REQUIRE(nx <= 4000 && ny <= 4000 && nz <= 4000 && nu <= 4000 && nv <= 4000,
                "geom3d: dimensions must be <= 4000");
    const float one[3] = {1.f, 1.f, 1.f};
    const float *vs = vol_scale ? vol_scale : one;
    AMX_REQUIRE(vs[0] > 0 && vs[1] > 0 && vs[2] > 0, "geom3d: voxel sizes must be positive");
    for (size_t k = 0; k < per * (size_t)na; ++k)
        AMX_REQUIRE(std::isfinite(vecs[k]), "geom3d: non-finite vector component (angle %zu)",
                    k / per);
    if (cyl)
        for (unsigned a = 0; a < na; ++a)
            AMX_REQUIRE(vecs[13 * (size_t)a + 12] > 0.0, "geom3d: cylinder radius must be positive "
                        "(angle %u)", a);

    auto *g = new Geom3D();
    g->kind = kind;
    g->nx = nx; g->ny = ny; g->nz = nz; g->na = na; g->nu = nu; g->nv = nv;
    for (int k = 0; k < 3; ++k) g->vol_scale[k] = vs[k];
    AMX_HIP(hipGetDevice(&g->device));
    g->vecs.assign(vecs, vecs + per * (size_t)na);
    g->fp_host.resize(na);
    g->dir_count.assign(3, 0);

    std::vector<BP3DAngle> bd(na), bf(na), b2(na);
    std::vector<CylBPAngle> bc(cyl ? na : 0);
    std::vector<float> wd(na, 0.f), w2(na, 0.f);
    for (unsigned a = 0; a < na; ++a) {
        const double *p = vecs + per * (size_t)a;
        int dir;
        if (cyl) {
            // cuda/3d/cone_cyl.cu:486-500: |src - detector centre| per axis
            // in world units, compared as float
            dir = principal_axis(std::fabs((float)(p[0] - p[3])), std::fabs((float)(p[1] - p[4])),
                                 std::fabs((float)(p[2] - p[5])));
            FP3DAngle &f = g->fp_host[a];
            f = FP3DAngle{};
            f.dir = dir;
            g->dir_count[dir]++;
            // rays in voxel units: every point scales by 1 / voxel size per
            // axis (the world-unit ray setup of cylcone_FP_t,
            // cuda/3d/cone_cyl.cu:260-284, is the same line)
            V3 ax, bx, cx;
            cyl_axes(p, ax, bx, cx);
            const double A3[3] = {ax.x, ax.y, ax.z}, B3[3] = {bx.x, bx.y, bx.z},
                         C3[3] = {cx.x, cx.y, cx.z};
            const int *pm = kPerm[dir];
            for (int k = 0; k < 3; ++k) {
                const int c = pm[k];
                const double iv = 1.0 / vs[c];
                f.s[k] = p[c] * iv;
                f.p0[k] = C3[c] * iv;
                f.u[k] = A3[c] * iv;
                f.b[k] = B3[c] * iv;
                f.v[k] = p[9 + c] * iv;
            }
            f.dth = std::sqrt(p[6] * p[6] + p[7] * p[7] + p[8] * p[8]) / p[12];
            cyl_bp_coeffs(p, bc[a]);
            continue;
        }
        if (kind == ASTRA_AMD_GEOM_CONE) {
            // cuda/3d/cone_fp.cu:378-387: |src - detector centre| per axis,
            // evaluated in double and compared as float
            float d3[3];
            for (int k = 0; k < 3; ++k)
                d3[k] = std::fabs((float)(p[k] - (p[3 + k] + nu * p[6 + k] * 0.5f +
                                                   nv * p[9 + k] * 0.5f)));
            dir = principal_axis(d3[0], d3[1], d3[2]);
        } else {
            // cuda/3d/par3d_fp.cu:478-487
            dir = principal_axis(std::fabs((float)p[0]), std::fabs((float)p[1]),
                                 std::fabs((float)p[2]));
        }
        FP3DAngle &f = g->fp_host[a];
        f = FP3DAngle{};
        f.dir = dir;
        g->dir_count[dir]++;
        const int *pm = kPerm[dir];
        for (int k = 0; k < 3; ++k) {
            int c = pm[k];
            f.s[k] = p[c];
            f.p0[k] = p[3 + c] + 0.5 * p[6 + c] + 0.5 * p[9 + c];
            f.u[k] = p[6 + c];
            f.v[k] = p[9 + c];
        }
        if (kind == ASTRA_AMD_GEOM_CONE) {
            cone_bp_coeffs(p, ASTRA_AMD_KER3D_DEFAULT, vs, bd[a]);
            cone_bp_coeffs(p, ASTRA_AMD_KER3D_FDK_WEIGHTING, vs, bf[a]);
            cone_bp_coeffs(p, ASTRA_AMD_KER3D_2D_WEIGHTING, vs, b2[a]);
        } else {
            par3d_bp_coeffs(p, ASTRA_AMD_KER3D_DEFAULT, vs, bd[a], wd[a]);
            par3d_bp_coeffs(p, ASTRA_AMD_KER3D_2D_WEIGHTING, vs, b2[a], w2[a]);
            bf[a] = bd[a];
        }
    }
    auto zconst = [&](const std::vector<BP3DAngle> &b) {
        for (unsigned a = 0; a < na; ++a)
            if (b[a].nu[2] != 0.f || b[a].dn[2] != 0.f) return false;
        return true;
    };
    g->zc_default = !cyl && zconst(bd);
    g->zc_fdk = !cyl && zconst(bf);
    g->zc_2d = !cyl && zconst(b2);
    try {
        g->fp = DevArray(sizeof(FP3DAngle) * na);
        g->fp.upload(g->fp_host.data(), sizeof(FP3DAngle) * na);
        g->bp_default = DevArray(sizeof(BP3DAngle) * na);
        g->bp_default.upload(bd.data(), sizeof(BP3DAngle) * na);
        g->bp_fdk = DevArray(sizeof(BP3DAngle) * na);
        g->bp_fdk.upload(bf.data(), sizeof(BP3DAngle) * na);
        g->bp_2d = DevArray(sizeof(BP3DAngle) * na);
        g->bp_2d.upload(b2.data(), sizeof(BP3DAngle) * na);
        g->bp_w_default = DevArray(sizeof(float) * na);
        g->bp_w_default.upload(wd.data(), sizeof(float) * na);
        g->bp_w_2d = DevArray(sizeof(float) * na);
        g->bp_w_2d.upload(w2.data(), sizeof(float) * na);
        if (cyl) {
            g->bp_cyl = DevArray(sizeof(CylBPAngle) * na);
            g->bp_cyl.upload(bc.data(), sizeof(CylBPAngle) * na);
        }
        for (int k = 0; k < 3; ++k) {
            std::vector<int> idx;
            for (unsigned a = 0; a < na; ++a)
                if (g->fp_host[a].dir == k) idx.push_back((int)a);
            if (!idx.empty()) {
                g->group_idx[k] = DevArray(sizeof(int) * idx.size());
                g->group_idx[k].upload(idx.data(), sizeof(int) * idx.size());
            }
        }
    } catch (...) {
        delete g;
        throw;
    }
    return g;
}

// ----------------------------------------------------------- 2D ----------
Geom2D *geom2d_build(int kind, unsigned w, unsigned h, unsigned na, unsigned nd,
                     const float *vecs) {
    AMX_REQUIRE(kind == ASTRA_AMD_GEOM_PAR || kind == ASTRA_AMD_GEOM_FAN,
                "geom2d: kind must be PAR or FAN (got %d)", kind);
    AMX_REQUIRE(w > 0 && h > 0 && na > 0 && nd > 0, "geom2d: all dimensions must be positive");
    AMX_REQUIRE(w <= 16000 && h <= 16000 && nd <= 1000000, "geom2d: dimensions too large");
    AMX_REQUIRE(vecs != nullptr, "geom2d: vecs is NULL");
    std::vector<FP2DAngle> fp(na);
    std::vector<BP2DAngle> bp(na), bpf(na);
    for (unsigned a = 0; a < na; ++a)
        angle2d_params(kind, vecs + 6 * (size_t)a, a, fp[a], bp[a], bpf[a]);

    auto *g = new Geom2D();
    g->kind = kind;
    g->w = w; g->h = h; g->na = na; g->nd = nd;
    g->vecs.assign(vecs, vecs + 6 * (size_t)na);
    try {
        AMX_HIP(hipGetDevice(&g->device));
        g->fp = DevArray(sizeof(FP2DAngle) * na);
        g->fp.upload(fp.data(), sizeof(FP2DAngle) * na);
        g->bp = DevArray(sizeof(BP2DAngle) * na);
        g->bp.upload(bp.data(), sizeof(BP2DAngle) * na);
        g->bp_fbp = DevArray(sizeof(BP2DAngle) * na);
        g->bp_fbp.upload(bpf.data(), sizeof(BP2DAngle) * na);
    } catch (...) {
        delete g;
        throw;
    }
    return g;
}

// Per-angle FP and BP coefficients of one 2D projection vector (normalised
// [rayX, rayY, dSX, dSY, uX, uY] or [srcX, srcY, ...]).
void angle2d_params(int kind, const float *p, unsigned a, FP2DAngle &f, BP2DAngle &bpa,
                    BP2DAngle &bpfa) {
    {
        for (int k = 0; k < 6; ++k)
            AMX_REQUIRE(std::isfinite(p[k]), "geom2d: non-finite vector component (angle %u)", a);
        f.p0[0] = (double)p[2] + 0.5 * (double)p[4];
        f.p0[1] = (double)p[3] + 0.5 * (double)p[5];
        f.u[0] = p[4];
        f.u[1] = p[5];
        f.r[0] = p[0];
        f.r[1] = p[1];
        for (int k = 0; k < 6; ++k) f.raw[k] = p[k];
        // cuda/2d/par_fp.cu:272 (par only; fan decides per ray)
        f.vertical = std::fabs(p[0]) <= std::fabs(p[1]) ? 1 : 0;
        f.pad_ = 0;
        if (kind == ASTRA_AMD_GEOM_PAR) {
            // cuda/2d/par_bp.cu:172-178: T = x c - y_row s + off with
            // y_row = -y_world, weight |r| / |d|
            double rx = p[0], ry = p[1], dsx = p[2], dsy = p[3], ux = p[4], uy = p[5];
            double d = ux * ry - uy * rx;
            AMX_REQUIRE(d != 0.0, "geom2d: detector parallel to rays (angle %u)", a);
            double c = ry / d, s = -rx / d, off = (dsy * rx - dsx * ry) / d;
            BP2DAngle &b = bpa;
            b.num[0] = (float)(off + 1.5);  // -0.5 texel centre, +2 padding
            b.num[1] = (float)c;
            b.num[2] = (float)s;
            b.den[0] = 1.f; b.den[1] = 0.f; b.den[2] = 0.f;
            b.w = (float)(std::sqrt(rx * rx + ry * ry) / std::fabs(d));
            b.pad_ = 0.f;
            bpfa = b;
        } else {
            // DevFanParams, cuda/2d/fan_bp.cu:240-271
            double sx = p[0], sy = p[1], dsx = p[2], dsy = p[3], ux = p[4], uy = p[5];
            double smdx = sx - dsx, smdy = sy - dsy;
            auto det2 = [](double ax, double ay, double bx, double by) { return ax * by - ay * bx; };
            for (int fbp = 0; fbp < 2; ++fbp) {
                double sc = fbp ? 1.0 / det2(ux, uy, sx, sy)
                                : std::sqrt(ux * ux + uy * uy) / det2(ux, uy, smdx, smdy);
                double nc = sc * det2(sx, sy, dsx, dsy), nxc = sc * smdy, nyc = -sc * smdx;
                double dc = fbp ? 1.0 : sc * det2(ux, uy, sx, sy), dxc = sc * uy, dyc = -sc * ux;
                BP2DAngle &b = fbp ? bpfa : bpa;
                b.num[0] = (float)(nc + 1.5 * dc);  // -0.5 texel centre, +2 padding
                b.num[1] = (float)(nxc + 1.5 * dxc);
                b.num[2] = (float)(nyc + 1.5 * dyc);
                b.den[0] = (float)dc;
                b.den[1] = (float)dxc;
                b.den[2] = (float)dyc;
                b.w = 1.f;
                b.pad_ = 0.f;
            }
        }
    }
}

// ------------------------------------------------------- filters ---------
namespace {
// Window of the reference's genFilter for bin k >= 1 of nb = n_fft/2 + 1
// bins (src/Filters.cpp:151-472); w = 2 pi k / n_fft.  Evaluated in double.
double filter_window(int type, unsigned k, unsigned nb, double w, double d, double param) {
    const double fk = (double)k, nm1 = (double)nb - 1.0;
    switch (type) {
    case ASTRA_AMD_FILTER_SHEPPLOGAN: {
        const double x = w / 2.0 / d;
        return std::sin(x) / x;
    }
    case ASTRA_AMD_FILTER_COSINE:
        return std::cos(w / 2.0 / d);
    case ASTRA_AMD_FILTER_HAMMING:
        return 0.54 + 0.46 * std::cos(w / d);
    case ASTRA_AMD_FILTER_HANN:
        return (1.0 + std::cos(w / d)) / 2.0;
    case ASTRA_AMD_FILTER_TUKEY: {
        const double alpha = param < 0 ? 0.5 : param;
        const double half_n = (double)nb / 2.0;
        const double enum_term = alpha * half_n, denom = (1.0 - alpha) * half_n;
        if (half_n - enum_term <= fk && fk <= half_n + enum_term) return 1.0;
        return 0.5 * (1.0 + std::cos(M_PI * (fk - enum_term) / denom));
    }
    case ASTRA_AMD_FILTER_LANCZOS: {
        const double x = M_PI * (2.0 * fk / nm1 - 1.0);
        return std::fabs(x) > 0.001 ? std::sin(x) / x : 1.0;
    }
    case ASTRA_AMD_FILTER_TRIANGULAR:
        return 2.0 / nm1 * (nm1 / 2.0 - std::fabs(fk - nm1 / 2.0));
    case ASTRA_AMD_FILTER_GAUSSIAN: {
        const double sigma = param < 0 ? 0.4 : param;
        const double q = nm1 / 2.0, t = (fk - q) / (sigma * q);
        return std::exp(-0.5 * t * t);
    }
    case ASTRA_AMD_FILTER_BARTLETTHANN:
        return 0.62 - 0.48 * std::fabs(fk / nm1 - 0.5) - 0.38 * std::cos(2.0 * M_PI * fk / nm1);
    case ASTRA_AMD_FILTER_BLACKMAN: {
        const double alpha = param < 0 ? 0.16 : param;
        return (1.0 - alpha) / 2.0 - 0.5 * std::cos(M_PI * fk / nm1) +
               alpha / 2.0 * std::cos(2.0 * M_PI * fk / nm1);
    }
    case ASTRA_AMD_FILTER_NUTTALL:
    case ASTRA_AMD_FILTER_BLACKMANHARRIS:
    case ASTRA_AMD_FILTER_BLACKMANNUTTALL:
    case ASTRA_AMD_FILTER_FLATTOP: {
        static const double a[4][5] = {{0.355768, 0.487396, 0.144232, 0.012604, 0.0},
                                       {0.35875, 0.48829, 0.14128, 0.01168, 0.0},
                                       {0.3635819, 0.4891775, 0.1365995, 0.0106411, 0.0},
                                       {1.0, 1.93, 1.29, 0.388, 0.032}};
        const double *c = a[type - ASTRA_AMD_FILTER_NUTTALL];
        const double b = M_PI * fk / nm1;
        return c[0] - c[1] * std::cos(2 * b) + c[2] * std::cos(4 * b) - c[3] * std::cos(6 * b) +
               c[4] * std::cos(8 * b);
    }
    case ASTRA_AMD_FILTER_KAISER: {
        // the reference evaluates the Bessel function J0 here (not I0)
        const double alpha = param < 0 ? 3.0 : param;
        const double x = 2.0 * fk / nm1 - 1.0;
        return ::j0(M_PI * alpha * std::sqrt(1.0 - x * x)) / ::j0(M_PI * alpha);
    }
    case ASTRA_AMD_FILTER_PARZEN: {
        const double q = fk / nm1;
        return q <= 0.5 ? 1.0 - 6.0 * q * q * (1.0 - q) : 2.0 * (1.0 - q) * (1.0 - q) * (1.0 - q);
    }
    default:
        return 1.0;  // Ram-Lak
    }
}
}  // namespace

void filter_half_spectrum(int type, unsigned n_det, double d, double param,
                          std::vector<double> &out) {
    AMX_REQUIRE(type >= ASTRA_AMD_FILTER_RAMLAK && type <= ASTRA_AMD_FILTER_PARZEN,
                "filter: unknown window type %d", type);
    AMX_REQUIRE(d > 0, "filter: FilterD must be positive");
    ramlak_half_spectrum(n_det, d, out);
    const unsigned n = next_pow2(2 * n_det), nb = n / 2 + 1;
    for (unsigned k = 1; k < nb; ++k) {
        const double w = 2.0 * M_PI * (double)k / (double)n;
        out[k] *= filter_window(type, k, nb, w, d, param);
    }
}

// In-place iterative radix-2 complex FFT (forward, unscaled) in double.
static void fft_double(std::vector<double> &re, std::vector<double> &im) {
    const size_t n = re.size();
    for (size_t i = 1, j = 0; i < n; ++i) {
        size_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            std::swap(re[i], re[j]);
            std::swap(im[i], im[j]);
        }
    }
    for (size_t len = 2; len <= n; len <<= 1) {
        const double ang = -2.0 * M_PI / (double)len;
        for (size_t i = 0; i < n; i += len)
            for (size_t k = 0; k < len / 2; ++k) {
                const double c = std::cos(ang * (double)k), sn = std::sin(ang * (double)k);
                const double xr = re[i + k + len / 2], xi = im[i + k + len / 2];
                const double tr = xr * c - xi * sn, ti = xr * sn + xi * c;
                re[i + k + len / 2] = re[i + k] - tr;
                im[i + k + len / 2] = im[i + k] - ti;
                re[i + k] += tr;
                im[i + k] += ti;
            }
    }
}

void custom_filter_spectra(unsigned n_det, unsigned n_rows, int real_space, const float *data,
                           unsigned width, std::vector<float> &out) {
    const unsigned n = next_pow2(2 * n_det), nb = n / 2 + 1;
    out.assign((size_t)n_rows * n * 2, 0.f);
    if (!real_space) {
        // Fourier-domain filter: nb real bins per row (cuda/2d/fft.cu:447-490);
        // the full spectrum is its even extension
        AMX_REQUIRE(width == nb, "filter: a Fourier-domain filter needs %u bins per row, got %u",
                    nb, width);
        for (unsigned r = 0; r < n_rows; ++r)
            for (unsigned k = 0; k < n; ++k)
                out[((size_t)r * n + k) * 2] = data[(size_t)r * width + (k <= n / 2 ? k : n - k)];
        return;
    }
    // real-space kernel centred at width/2, wrapped into the padded length
    // and transformed (cuda/2d/fft.cu:491-595)
    AMX_REQUIRE(width > 0, "filter: empty real-space filter");
    const unsigned used = std::min(width, n), start = (width - used) / 2, shift = width / 2;
    std::vector<double> re(n), im(n);
    for (unsigned r = 0; r < n_rows; ++r) {
        std::fill(re.begin(), re.end(), 0.0);
        std::fill(im.begin(), im.end(), 0.0);
        for (unsigned j = start; j < start + used; ++j)
            re[(j + n - shift) % n] = data[(size_t)r * width + j];
        fft_double(re, im);
        for (unsigned k = 0; k < n; ++k) {
            out[((size_t)r * n + k) * 2] = (float)re[k];
            out[((size_t)r * n + k) * 2 + 1] = (float)im[k];
        }
    }
}

void parker_weights(const std::vector<double> &angles, unsigned nu, double sod, double odd,
                    double du, std::vector<float> &out) {
    // FDK_PreWeight short-scan branch + devFDK_ParkerWeight
    // (cuda/3d/fdk.cu:115-163, 201-251): angles relative to the start of
    // the scanned range, wrapped into [0, 2 pi); weight of ray (beta, gamma)
    // from the central fan's Radon-space position, times range / pi.
    const size_t na = angles.size();
    out.assign(na * nu, 1.f);
    if (na < 2) return;
    double da = angles[1] - angles[0];
    while (da < -M_PI) da += 2 * M_PI;
    while (da >= M_PI) da -= 2 * M_PI;
    const double base = da >= 0 ? angles[0] : angles[na - 1];
    std::vector<double> rel(na);
    for (size_t i = 0; i < na; ++i) {
        double f = angles[i] - base;
        while (f >= 2 * M_PI) f -= 2 * M_PI;
        while (f < 0) f += 2 * M_PI;
        rel[i] = f;
    }
    const double range = std::fabs(rel[na - 1] - rel[0]) / (double)(na - 1) * (double)na;
    const double scale = range / M_PI;
    const double D = sod + odd;
    const double fan = std::fabs(std::atan(du * (nu * 0.5) / D));
    for (size_t i = 0; i < na; ++i) {
        const double beta = rel[i];
        for (unsigned u = 0; u < nu; ++u) {
            const double gamma = std::atan((u - 0.5 * nu + 0.5) * du / D);
            double w;
            if (beta <= 0.0) {
                w = 0.0;
            } else if (beta <= 2.0 * (fan + gamma)) {
                w = std::sin((M_PI / 4.0) * beta / (fan + gamma));
                w *= w;
            } else if (beta <= M_PI + 2 * gamma) {
                w = 1.0;
            } else if (beta <= M_PI + 2 * fan) {
                w = std::sin((M_PI / 4.0) * (M_PI + 2.0 * fan - beta) / (fan - gamma));
                w *= w;
            } else {
                w = 0.0;
            }
            out[i * nu + u] = (float)(w * scale);
        }
    }
}

// In-place radix-2 FFT of a power-of-two length (FP64, host; forward sign).
static void host_fft(std::vector<std::complex<double>> &a) {
    const size_t n = a.size();
    for (size_t i = 1, j = 0; i < n; ++i) {  // bit-reversal permutation
        size_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(a[i], a[j]);
    }
    for (size_t len = 2; len <= n; len <<= 1) {
        const double ang = -2.0 * M_PI / (double)len;
        for (size_t i = 0; i < n; i += len) {
            for (size_t k = 0; k < len / 2; ++k) {
                const std::complex<double> w(std::cos(ang * (double)k), std::sin(ang * (double)k));
                const std::complex<double> u = a[i + k], t = a[i + k + len / 2] * w;
                a[i + k] = u + t;
                a[i + k + len / 2] = u - t;
            }
        }
    }
}

void ramlak_half_spectrum(unsigned n_det, double d, std::vector<double> &out) {
    // genFilter(FILTER_RAMLAK) (src/Filters.cpp:133-160, 473-485) with the
    // spatial kernel of generateRampFilter (src/Filters.cpp:41-68):
    // h[0] = 1/4, h[j odd] = -1/(pi min(j, N-j))^2, h[j even] = 0.  h is real
    // and even, so its DFT is real; Filter = 2 Re DFT(h)[k], k = 0 .. N/2,
    // zeroed where 2 pi k / N > pi d.  The DFT is an FP64 FFT.
    unsigned n = next_pow2(2 * n_det);
    unsigned nb = n / 2 + 1;
    std::vector<std::complex<double>> h(n, 0.0);
    h[0] = 0.25;
    for (unsigned j = 1; j < n; j += 2) {
        const double jj = (2 * j > n) ? (double)(n - j) : (double)j;
        h[j] = -1.0 / ((M_PI * jj) * (M_PI * jj));
    }
    host_fft(h);
    out.assign(nb, 0.0);
    for (unsigned k = 0; k < nb; ++k) {
        const double wv = 2.0 * M_PI * (double)k / (double)n;
        out[k] = (wv > M_PI * d) ? 0.0 : 2.0 * h[k].real();
    }
}

}  // namespace amx
