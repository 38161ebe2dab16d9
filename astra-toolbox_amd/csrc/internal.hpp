// Internal declarations shared by the host and device sources of
// libastra_amd.  Nothing here is part of the C ABI (include/astra_amd.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/astra_amd.h"

namespace amx {

// ------------------------------------------------------------- errors ----
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] void fail(int code, const char *fmt, ...);

#define AMX_HIP(call)                                                                    \
    do {                                                                                 \
        hipError_t amx_e_ = (call);                                                      \
        if (amx_e_ != hipSuccess)                                                        \
            ::amx::fail(ASTRA_AMD_ERR_HIP, "%s failed: %s (%s:%d)", #call,               \
                        hipGetErrorString(amx_e_), __FILE__, __LINE__);                  \
    } while (0)

#define AMX_REQUIRE(cond, ...)                                                           \
    do {                                                                                 \
        if (!(cond)) ::amx::fail(ASTRA_AMD_ERR_ARG, __VA_ARGS__);                        \
    } while (0)

// Checks the launch status of the kernel just enqueued.
#define AMX_LAUNCHED() AMX_HIP(hipGetLastError())

// ----------------------------------------------------------- scratch -----
// Stream-ordered device scratch buffer (hipMallocAsync / hipFreeAsync on the
// stream of the call that uses it).  The default memory pool of every device
// is configured to keep freed blocks, so repeated calls do not go back to
// the driver.
class Scratch {
  public:
    Scratch(size_t bytes, hipStream_t s);
    ~Scratch();
    Scratch(const Scratch &) = delete;
    Scratch &operator=(const Scratch &) = delete;
    template <class T> T *as() const { return static_cast<T *>(p_); }

  private:
    void *p_ = nullptr;
    hipStream_t s_;
};

void configure_pool_once();
// Bytes of the current device's stream-ordered pool in use now and at the
// high-water mark (optionally reset to the current use).
void scratch_usage(size_t &in_use, size_t &peak, bool reset_peak);

// Persistent device allocation owned by a handle (hipMalloc / hipFree).
class DevArray {
  public:
    DevArray() = default;
    explicit DevArray(size_t bytes);
    ~DevArray();
    DevArray(const DevArray &) = delete;
    DevArray &operator=(const DevArray &) = delete;
    DevArray(DevArray &&o) noexcept : p_(o.p_), bytes_(o.bytes_) { o.p_ = nullptr; o.bytes_ = 0; }
    DevArray &operator=(DevArray &&o) noexcept;
    void upload(const void *host, size_t bytes);
    template <class T> T *as() const { return static_cast<T *>(p_); }
    size_t bytes() const { return bytes_; }

  private:
    void *p_ = nullptr;
    size_t bytes_ = 0;
};

// ---------------------------------------------------- device params ------
// Per-angle forward-projection ray data, FP64, already permuted into the
// (c0, c1, c2) = (principal axis, in-slice fast axis, in-slice slow axis)
// order of the angle's direction group.
// Cylindrical detector (cyl_cone_vec): s = source, p0 = cylinder centre,
// u = radial vector (axis -> detector centre, length R), b = tangential
// vector (length R), v = one pixel step along the axis, dth = angular step
// per pixel; pixel (iu, iv) sits at p0 + u cos(th) + b sin(th) +
// v (iv - nv/2 + 0.5) with th = (iu - nu/2 + 0.5) dth.
struct FP3DAngle {
    double s[3];   // source (cone, cylinder) or ray direction (par3d)
    double p0[3];  // centre of detector pixel (0, 0) (cylinder: axis centre)
    double u[3];   // one detector pixel step along u (cylinder: radial)
    double v[3];   // one detector pixel step along v
    double b[3];   // cylinder: tangential
    double dth;    // cylinder: angular step per pixel
    int dir;       // 0 = x, 1 = y, 2 = z principal axis
    int pad_;
};

// Per-angle cylindrical-detector BP coefficients (float32, world units of
// the translated frame), DevCylConeParamsBP of cuda/3d/cone_cyl_bp.cu:
// 165-230: rows (x, y, z, w) of the source-relative ray in the (radial,
// tangential, axial) frame scaled by sqrt(|u x v|); the source in that
// frame; c = |s - centre|^2 in the radial-tangential plane - R^2.
struct CylBPAngle {
    float ca[4], cb[4], cc[4];
    float sa, sb, sc, c;
    float inv_dth, inv_nv, inv_r, pad_;
};

// Per-angle voxel-driven back-projection coefficients (float32), in padded
// texel coordinates: p_u = (nu . [x,y,z,1]) / (dn . [x,y,z,1]) = U - 0.5 + 2,
// likewise p_v; for par3d dn = (0,0,0,1).
struct BP3DAngle {
    float nu[4];
    float nv[4];
    float dn[4];
};

struct FP2DAngle {
    double p0[2];  // centre of detector pixel 0
    double u[2];   // detector pixel step
    double r[2];   // ray direction (par) or source position (fan)
    float raw[6];  // the float32 vector as given (fan orientation test)
    int vertical;  // par: per-angle orientation (1 = rows are slices)
    int pad_;
};

struct BP2DAngle {
    float num[3];  // T + 1.5 numerator: c + x*cx + y*cy (y is world y)
    float den[3];  // denominator (par: 1, 0, 0)
    float w;       // par: per-angle weight |r|/|d|
    float pad_;
};

// ------------------------------------------------------ handles ----------
struct Filter {
    unsigned n_det, n_fft;
    int device;
    bool identity = false;  // FilterType 'none': rows are left unfiltered
    // general = false: one real, even spectrum for every row, float[n_fft].
    // general = true: n_spec complex spectra float2[n_spec][n_fft] (row r
    // uses spectrum r % n_spec).
    bool general = false;
    unsigned n_spec = 1;
    DevArray spec;
    DevArray twiddle;  // float2[n_fft]: exp(-2 pi i k / n_fft)
};

struct Geom3D {
    int kind;  // ASTRA_AMD_GEOM_PAR3D / _CONE / _CYLCONE
    unsigned nx, ny, nz, na, nu, nv;
    float vol_scale[3];
    int device;
    std::vector<double> vecs;  // normalised host copy, 12 per angle
    std::vector<FP3DAngle> fp_host;
    std::vector<int> dir_count;  // angles per direction group
    DevArray fp;                 // FP3DAngle[na]
    DevArray bp_default, bp_fdk, bp_2d;  // BP3DAngle[na]
    DevArray bp_w_default, bp_w_2d;      // float[na] par3d weights
    DevArray bp_cyl;                     // CylBPAngle[na] (cylindrical detector)
    // per BP coefficient set: U and the denominator do not depend on z for
    // any angle (nu[2] == dn[2] == 0, e.g. circular cone with a vertical
    // detector), so the BP hoists them out of its z loop
    bool zc_default = false, zc_fdk = false, zc_2d = false;
    DevArray group_idx[3];               // int[dir_count[k]] angle lists
    // FDK pre-weights [nv][nu] and default Ram-Lak filter, built on first use
    mutable std::mutex fdk_mu;
    mutable DevArray fdk_w;
    mutable DevArray fdk_parker;  // [na][nu] short-scan weights, built on first use
    mutable std::unique_ptr<Filter> fdk_filter;
};

struct Geom2D {
    int kind;  // ASTRA_AMD_GEOM_PAR / _FAN
    unsigned w, h, na, nd;
    int device;
    std::vector<float> vecs;  // 6 per angle
    DevArray fp;              // FP2DAngle[na]
    DevArray bp, bp_fbp;      // BP2DAngle[na]
    // FBP state built on first use: the default Ram-Lak filter, the fan
    // pre-weights [nd] and short-scan Parker weights [na][nd]
    mutable std::mutex fbp_mu;
    mutable std::unique_ptr<Filter> fbp_ramlak;
    mutable DevArray fbp_w;
    mutable DevArray fbp_parker;
};

// ------------------------------------------------------ profiling --------
struct ProfToken {
    long index = -1;
    bool active = false;
};
bool prof_on();
ProfToken prof_begin(const char *name, double units, hipStream_t s);
void prof_end(const ProfToken &t, hipStream_t s);
void prof_enable(bool on);
bool prof_read(const char *name, double &ms, double &units, unsigned long long &launches);
// 16-byte gather calibration (calib.hip): lane-loads per second.
double gather_calibrate(int mode, unsigned iters, hipStream_t s);

// ------------------------------------------------------ launchers --------
// 3D
// mode: OUT_ADD / OUT_SET / OUT_RESID (out = (b + scale*Ax) * w); padded_out
// writes the padded angle-major layout [a][v+4][u+4] consumed by the BP.
// ss: rays per detector pixel and dimension (FP) / sub-voxels per voxel
// dimension (BP) -- DetectorSuperSampling / VoxelSuperSampling
void fp3d_launch(const Geom3D &g, const float *vol, float *proj, float out_scale, int mode,
                 const float *resid_b, const float *resid_w, bool padded_out, hipStream_t s,
                 unsigned ss = 1);
// BP from user-layout projections (makes the padded copy internally)
void bp3d_launch(const Geom3D &g, const float *proj, float *vol, float out_scale, int kernel,
                 int mode, const float *upd_w, float lo, float hi, hipStream_t s,
                 unsigned ss = 1);
// BP from an existing padded angle-major copy
void bp3d_launch_padded(const Geom3D &g, const float *pp, float *vol, float out_scale, int kernel,
                        int mode, const float *upd_w, float lo, float hi, hipStream_t s,
                        unsigned ss = 1);
void proj_to_padded_am(const Geom3D &g, const float *proj, float *pp, hipStream_t s);
// par3d FP with the sum_square_weights kernel (squared weights, no volume)
void fp3d_sumsq_launch(const Geom3D &g, float *proj, float out_scale, int mode, hipStream_t s);
size_t padded_am_elems(const Geom3D &g);
void bp3d_matched_launch(const Geom3D &g, const float *proj, float *vol, float out_scale,
                         int mode, hipStream_t s, unsigned ss = 1);
// 2D
void fp2d_launch(const Geom2D &g, const float *vol, float *sino, float out_scale, int mode,
                 const float *resid_b, const float *resid_w, bool padded_out, unsigned rays,
                 hipStream_t s);
// FP of angles [a0, a0 + na_sub) only (rows a0.. of the output / residual)
void fp2d_launch_range(const Geom2D &g, const float *vol, float *sino, float out_scale, int mode,
                       const float *resid_b, const float *resid_w, bool padded_out, unsigned rays,
                       unsigned a0, unsigned na_sub, hipStream_t s);
bool bp2d_lds();
bool fp2d_pairs();
bool bp2d_pairs();
size_t fp2d_copies_elems(const Geom2D &g);
void fp2d_run(const Geom2D &g, const float *vol, float *sino, float out_scale, int mode,
              const float *resid_b, const float *resid_w, bool padded_out, unsigned rays,
              unsigned a0, unsigned na_sub, float *copies, bool prof, hipStream_t s);
// FP with the volume copies in a fixed per-GPU array (no allocation)
void fp2d_launch_noalloc(const Geom2D &g, const float *vol, float *sino, float out_scale,
                         int mode, unsigned rays, hipStream_t s);
size_t fp2d_noalloc_max_floats();
// SART's unweighted BP of one padded sinogram row: vol = clamp(vol + w * BP)
void bp2d_sart_launch(const Geom2D &g, const float *sp_row, float *vol, unsigned angle,
                      float out_scale, const float *upd_w, float lo, float hi, hipStream_t s);
void bp2d_launch(const Geom2D &g, const float *sino, float *vol, float out_scale, bool fbp,
                 int mode, const float *upd_w, float lo, float hi, hipStream_t s,
                 unsigned ss = 1);
void bp2d_launch_padded(const Geom2D &g, const float *sp, float *vol, float out_scale, bool fbp,
                        int mode, const float *upd_w, float lo, float hi, hipStream_t s,
                        unsigned ss = 1);
void sino_to_padded(const Geom2D &g, const float *sino, float *sp, hipStream_t s);
size_t padded_sino_elems(const Geom2D &g);
void bp2d_matched_launch(const Geom2D &g, const float *sino, float *vol, float out_scale,
                         int mode, hipStream_t s, unsigned rays = 1);
// filtering
void filter_rows_launch(const Filter &f, float *rows, size_t n_rows, const float *weights,
                        size_t rows_per_weight, hipStream_t s, const float *weights2 = nullptr,
                        unsigned w2_rows = 1);
void fdk_weights(const Geom3D &g, std::vector<float> &w);  // [nv][nu] host
void fdk_parker(const Geom3D &g, std::vector<float> &w);  // [na][nu] host
// elementwise
void fill_launch(float *p, size_t n, float v, hipStream_t s);
void invert_launch(float *p, size_t n, float scale, hipStream_t s);
void mul_launch(float *p, const float *q, size_t n, hipStream_t s);
void copy_launch(float *dst, const float *src, size_t n, hipStream_t s);
void clamp_launch(float *p, size_t n, float lo, float hi, const float *lo_mask,
                  const float *hi_mask, hipStream_t s);
// rows[r][i] *= w[r / rows_per_weight][i]
void rowscale_launch(float *rows, size_t n_rows, unsigned len, const float *w,
                     size_t rows_per_weight, hipStream_t s);
// CGLS helpers: scalars live in device memory (double).
void dot_launch(const float *a, const float *b, size_t n, double *out, hipStream_t s);
void axpy_ratio_launch(float *y, const float *x, size_t n, const double *num, const double *den,
                       double sign, hipStream_t s);
void xpby_ratio_launch(float *p, const float *z, size_t n, const double *num, const double *den,
                       hipStream_t s);

// geometry helpers
Geom3D *geom3d_build(int kind, unsigned nx, unsigned ny, unsigned nz, unsigned na, unsigned nu,
                     unsigned nv, const double *vecs, const float *vol_scale);
Geom2D *geom2d_build(int kind, unsigned w, unsigned h, unsigned na, unsigned nd,
                     const float *vecs);
void angle2d_params(int kind, const float *p, unsigned a, FP2DAngle &f, BP2DAngle &bp,
                    BP2DAngle &bp_fbp);
void ramlak_half_spectrum(unsigned n_det, double d, std::vector<double> &out);
void filter_half_spectrum(int type, unsigned n_det, double d, double param,
                          std::vector<double> &out);
// short-scan Parker weights [angle][u]
void parker_weights(const std::vector<double> &angles, unsigned nu, double sod, double odd,
                    double du, std::vector<float> &out);
// complex spectra float2[n_rows][n_fft] (interleaved) of user filters
void custom_filter_spectra(unsigned n_det, unsigned n_rows, int real_space, const float *data,
                           unsigned width, std::vector<float> &out);
Filter *filter_build_custom(unsigned n_det, unsigned n_rows, int real_space, const float *data,
                            unsigned width);
unsigned next_pow2(unsigned n);

}  // namespace amx
