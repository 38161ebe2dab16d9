// extern "C" entry points of libastra_amd (declared in include/astra_amd.h).
// Every entry point converts exceptions into a status code and a
// thread-local message, the role CLogger::getLastErrMsg plays in the
// reference (include/astra/Logging.h:169).
#include <cmath>
#include <cstring>
#include <memory>
#include <string>

#include "devutil.hpp"
#include "internal.hpp"

namespace amx {
struct AlgoBase;
AlgoBase *algo_build(int type, Geom2D *g2, Geom3D *g3, float out_scale, int kernel,
                     unsigned rays, unsigned vss);
void algo_destroy(AlgoBase *a);
void algo_set_option(AlgoBase *a, const char *key, double v);
void algo_set_mask(AlgoBase *a, const char *key, const float *m);
void algo_set_order(AlgoBase *a, const int *order, unsigned n);
void algo_run(AlgoBase *a, const float *b, float *x, int iters, hipStream_t s);
double algo_residual(AlgoBase *a, const float *b, const float *x, hipStream_t s);
void set_abort_hook(astra_amd_abort_hook fn, void *user);
Filter *filter_build(unsigned n_det, const float *half, unsigned n_bins);
void fdk_run(const Geom3D &g, float *proj, float *vol, float out_scale, const Filter *f,
             bool short_scan, int mode, hipStream_t s);
void fdk_preweight_only(const Geom3D &g, float *proj, hipStream_t s);
void fdk_filter_run(const Geom3D &g, float *proj, const Filter *f, bool short_scan,
                    hipStream_t s);
void fbp2d_run(const Geom2D &g, float *sino, float *vol, float out_scale, const Filter *f,
               bool short_scan, int mode, hipStream_t s);
struct Batch2D;
Batch2D *batch2d_build(unsigned n, const int *kinds, const unsigned *ws, const unsigned *hs,
                       const unsigned *nas, const unsigned *nds, const float *vecs,
                       const float *oscs);
void batch2d_destroy(Batch2D *b);
void batch2d_offsets(const Batch2D &b, size_t *vol_off, size_t *sino_off);
void batch2d_sirt(Batch2D &b, const float *sinos, float *vols, int iters, double relaxation,
                  float lo, float hi, int run_mode, hipStream_t s);
void batch2d_fp(Batch2D &b, const float *vols, float *sinos, hipStream_t s);
}  // namespace amx

struct astra_amd_batch2d_s {
    amx::Batch2D *b;
    int device;
};

struct astra_amd_geom2d_s {
    amx::Geom2D *g;
};
struct astra_amd_geom3d_s {
    amx::Geom3D *g;
};
struct astra_amd_filter_s {
    amx::Filter *f;
};
struct astra_amd_algo_s {
    amx::AlgoBase *a;
    int device;
};

namespace {

thread_local std::string t_err;

template <class F> int guarded(F &&fn) {
    try {
        fn();
        return ASTRA_AMD_OK;
    } catch (const amx::Error &e) {
        t_err = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        t_err = "host allocation failed";
        return ASTRA_AMD_ERR_ALLOC;
    } catch (const std::exception &e) {
        t_err = e.what();
        return ASTRA_AMD_ERR_INTERNAL;
    } catch (...) {
        t_err = "unknown error";
        return ASTRA_AMD_ERR_INTERNAL;
    }
}

void check_device(int dev) {
    int cur = -1;
    AMX_HIP(hipGetDevice(&cur));
    AMX_REQUIRE(cur == dev, "handle belongs to device %d but device %d is current", dev, cur);
}

int out_mode(int mode) {
    AMX_REQUIRE(mode == ASTRA_AMD_MODE_ADD || mode == ASTRA_AMD_MODE_SET,
                "mode must be ASTRA_AMD_MODE_ADD or ASTRA_AMD_MODE_SET (got %d)", mode);
    return mode == ASTRA_AMD_MODE_ADD ? amx::OUT_ADD : amx::OUT_SET;
}

hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// A pitched device array (`rows` rows of `width` floats, rows `pitch` floats
// apart) seen as a contiguous one: the array itself when pitch == width,
// else a stream-ordered scratch copy (filled from the array when `load`),
// written back by store().
class Contig {
  public:
    Contig(const float *p, size_t width, size_t rows, size_t pitch, bool load, hipStream_t s)
        : src_(const_cast<float *>(p)), w_(width), rows_(rows), pitch_(pitch), s_(s) {
        AMX_REQUIRE(pitch >= width, "pitch (%zu floats) is smaller than the row width (%zu)",
                    pitch, width);
        if (pitch == width) {
            ptr_ = src_;
            return;
        }
        scr_.reset(new amx::Scratch(width * rows * sizeof(float), s));
        ptr_ = scr_->as<float>();
        if (load)
            AMX_HIP(hipMemcpy2DAsync(ptr_, w_ * sizeof(float), src_, pitch_ * sizeof(float),
                                     w_ * sizeof(float), rows_, hipMemcpyDeviceToDevice, s_));
    }
    float *get() const { return ptr_; }
    void store() {
        if (ptr_ != src_)
            AMX_HIP(hipMemcpy2DAsync(src_, pitch_ * sizeof(float), ptr_, w_ * sizeof(float),
                                     w_ * sizeof(float), rows_, hipMemcpyDeviceToDevice, s_));
    }

  private:
    float *src_, *ptr_ = nullptr;
    size_t w_, rows_, pitch_;
    hipStream_t s_;
    std::unique_ptr<amx::Scratch> scr_;
};

}  // namespace

extern "C" {

const char *astra_amd_last_error(void) { return t_err.c_str(); }

const char *astra_amd_version(void) { return "astra_amd 0.1 (gfx950)"; }

int astra_amd_set_device(int index) {
    return guarded([&] { AMX_HIP(hipSetDevice(index)); });
}

int astra_amd_stream_sync(void *stream) {
    return guarded([&] { AMX_HIP(hipStreamSynchronize(as_stream(stream))); });
}

// ------------------------------------------------------------- 2D -------
int astra_amd_geom2d_create(int kind, unsigned int vol_width, unsigned int vol_height,
                            unsigned int n_angles, unsigned int n_dets, const float *vecs,
                            astra_amd_geom2d *out) {
    return guarded([&] {
        AMX_REQUIRE(out != nullptr, "geom2d_create: out is NULL");
        *out = nullptr;
        amx::Geom2D *g = amx::geom2d_build(kind, vol_width, vol_height, n_angles, n_dets, vecs);
        *out = new astra_amd_geom2d_s{g};
    });
}

int astra_amd_geom2d_destroy(astra_amd_geom2d g) {
    return guarded([&] {
        if (!g) return;
        delete g->g;
        delete g;
    });
}

static void check2d(astra_amd_geom2d g, const astra_amd_params2d *p) {
    AMX_REQUIRE(g && g->g, "geometry handle is NULL");
    AMX_REQUIRE(p != nullptr, "params is NULL");
    check_device(g->g->device);
}

int astra_amd_fp2d(astra_amd_geom2d g, const float *vol, float *sino,
                   const astra_amd_params2d *p, int mode, void *stream) {
    return guarded([&] {
        check2d(g, p);
        AMX_REQUIRE(vol && sino, "fp2d: NULL data pointer");
        unsigned rays = p->rays_per_det ? p->rays_per_det : 1;
        amx::fp2d_launch(*g->g, vol, sino, p->output_scale, out_mode(mode), nullptr, nullptr,
                         false, rays, as_stream(stream));
    });
}

int astra_amd_fp2d_noalloc(astra_amd_geom2d g, const float *vol, float *sino,
                           const astra_amd_params2d *p, int mode, void *stream) {
    return guarded([&] {
        check2d(g, p);
        AMX_REQUIRE(vol && sino, "fp2d_noalloc: NULL data pointer");
        unsigned rays = p->rays_per_det ? p->rays_per_det : 1;
        amx::fp2d_launch_noalloc(*g->g, vol, sino, p->output_scale, out_mode(mode), rays,
                                 as_stream(stream));
    });
}

int astra_amd_bp2d(astra_amd_geom2d g, const float *sino, float *vol,
                   const astra_amd_params2d *p, int mode, void *stream) {
    return guarded([&] {
        check2d(g, p);
        AMX_REQUIRE(vol && sino, "bp2d: NULL data pointer");
        amx::bp2d_launch(*g->g, sino, vol, p->output_scale, false, out_mode(mode), nullptr,
                         -INFINITY, INFINITY, as_stream(stream),
                         p->rays_per_pixel_dim ? p->rays_per_pixel_dim : 1);
    });
}

int astra_amd_bp2d_fbp_weighted(astra_amd_geom2d g, const float *sino, float *vol,
                                const astra_amd_params2d *p, int mode, void *stream) {
    return guarded([&] {
        check2d(g, p);
        AMX_REQUIRE(vol && sino, "bp2d_fbp_weighted: NULL data pointer");
        AMX_REQUIRE(g->g->kind == ASTRA_AMD_GEOM_FAN, "bp2d_fbp_weighted: FAN geometry required");
        amx::bp2d_launch(*g->g, sino, vol, p->output_scale, true, out_mode(mode), nullptr,
                         -INFINITY, INFINITY, as_stream(stream));
    });
}

int astra_amd_bp2d_matched(astra_amd_geom2d g, const float *sino, float *vol,
                           const astra_amd_params2d *p, int mode, void *stream) {
    return guarded([&] {
        check2d(g, p);
        AMX_REQUIRE(vol && sino, "bp2d_matched: NULL data pointer");
        amx::bp2d_matched_launch(*g->g, sino, vol, p->output_scale, out_mode(mode),
                                 as_stream(stream), p->rays_per_det ? p->rays_per_det : 1);
    });
}

// ------------------------------------------------------------- 3D -------
int astra_amd_geom3d_create(int kind, unsigned int vol_x, unsigned int vol_y,
                            unsigned int vol_z, unsigned int n_angles, unsigned int det_u,
                            unsigned int det_v, const double *vecs, const float *vol_scale,
                            astra_amd_geom3d *out) {
    return guarded([&] {
        AMX_REQUIRE(out != nullptr, "geom3d_create: out is NULL");
        *out = nullptr;
        amx::Geom3D *g = amx::geom3d_build(kind, vol_x, vol_y, vol_z, n_angles, det_u, det_v,
                                           vecs, vol_scale);
        *out = new astra_amd_geom3d_s{g};
    });
}

int astra_amd_geom3d_destroy(astra_amd_geom3d g) {
    return guarded([&] {
        if (!g) return;
        delete g->g;
        delete g;
    });
}

static void check3d(astra_amd_geom3d g, const astra_amd_params3d *p) {
    AMX_REQUIRE(g && g->g, "geometry handle is NULL");
    AMX_REQUIRE(p != nullptr, "params is NULL");
    check_device(g->g->device);
}

static unsigned ss_of(unsigned v) { return v ? v : 1; }

int astra_amd_fp3d(astra_amd_geom3d g, const float *vol, float *proj,
                   const astra_amd_params3d *p, int mode, void *stream) {
    return guarded([&] {
        check3d(g, p);
        AMX_REQUIRE(vol && proj, "fp3d: NULL data pointer");
        AMX_REQUIRE(p->kernel >= 0 && p->kernel <= ASTRA_AMD_KER3D_MATCHED_BP,
                    "fp3d: unknown kernel %d", p->kernel);
        // cuda/3d/mem3d.cu:385-393: default and matched_bp only
        if (g->g->kind == ASTRA_AMD_GEOM_CYLCONE && p->kernel != ASTRA_AMD_KER3D_DEFAULT &&
            p->kernel != ASTRA_AMD_KER3D_MATCHED_BP)
            amx::fail(ASTRA_AMD_ERR_UNSUPPORTED, "fp3d: the cylindrical-detector geometry supports "
                      "the default and matched_bp kernels only");
        if (p->kernel == ASTRA_AMD_KER3D_SUM_SQUARE_WEIGHTS) {
            if (ss_of(p->rays_per_det_dim) > 1)
                amx::fail(ASTRA_AMD_ERR_UNSUPPORTED,
                          "fp3d: sum_square_weights with DetectorSuperSampling > 1 is not built");
            amx::fp3d_sumsq_launch(*g->g, proj, p->output_scale, out_mode(mode), as_stream(stream));
            return;
        }
        amx::fp3d_launch(*g->g, vol, proj, p->output_scale, out_mode(mode), nullptr, nullptr,
                         false, as_stream(stream), ss_of(p->rays_per_det_dim));
    });
}

int astra_amd_bp3d(astra_amd_geom3d g, const float *proj, float *vol,
                   const astra_amd_params3d *p, int mode, void *stream) {
    return guarded([&] {
        check3d(g, p);
        AMX_REQUIRE(vol && proj, "bp3d: NULL data pointer");
        const int m = out_mode(mode);
        switch (p->kernel) {
            case ASTRA_AMD_KER3D_DEFAULT:
            case ASTRA_AMD_KER3D_2D_WEIGHTING:
                amx::bp3d_launch(*g->g, proj, vol, p->output_scale, p->kernel, m, nullptr,
                                 -INFINITY, INFINITY, as_stream(stream),
                                 ss_of(p->rays_per_voxel_dim));
                break;
            case ASTRA_AMD_KER3D_FDK_WEIGHTING:
                AMX_REQUIRE(g->g->kind == ASTRA_AMD_GEOM_CONE,
                            "bp3d: fdk_weighting requires a CONE geometry");
                amx::bp3d_launch(*g->g, proj, vol, p->output_scale, p->kernel, m, nullptr,
                                 -INFINITY, INFINITY, as_stream(stream));
                break;
            case ASTRA_AMD_KER3D_MATCHED_BP:
                amx::bp3d_matched_launch(*g->g, proj, vol, p->output_scale, m, as_stream(stream),
                                         ss_of(p->rays_per_det_dim));
                break;
            case ASTRA_AMD_KER3D_SUM_SQUARE_WEIGHTS:
                amx::fail(ASTRA_AMD_ERR_UNSUPPORTED, "bp3d: sum_square_weights is FP-only");
            default:
                amx::fail(ASTRA_AMD_ERR_ARG, "bp3d: unknown kernel %d", p->kernel);
        }
    });
}

// ---------------------------------------------------------- filters -----
int astra_amd_ramlak_spectrum(unsigned int n_det, float d, float *out, unsigned int n_out) {
    return guarded([&] {
        AMX_REQUIRE(n_det > 0 && out != nullptr, "ramlak_spectrum: invalid arguments");
        std::vector<double> h;
        amx::ramlak_half_spectrum(n_det, d, h);
        AMX_REQUIRE(n_out == h.size(), "ramlak_spectrum: expected n_out = %zu", h.size());
        for (size_t k = 0; k < h.size(); ++k) out[k] = (float)h[k];
    });
}

int astra_amd_filter_spectrum(int type, unsigned int n_det, float d, float param, float *out,
                              unsigned int n_out) {
    return guarded([&] {
        AMX_REQUIRE(n_det > 0 && out != nullptr, "filter_spectrum: invalid arguments");
        std::vector<double> h;
        amx::filter_half_spectrum(type, n_det, d, param, h);
        AMX_REQUIRE(n_out == h.size(), "filter_spectrum: expected n_out = %zu", h.size());
        for (size_t k = 0; k < h.size(); ++k) out[k] = (float)h[k];
    });
}

int astra_amd_filter_create_custom(unsigned int n_det, unsigned int n_rows, int real_space,
                                   const float *data, unsigned int width,
                                   astra_amd_filter *out) {
    return guarded([&] {
        AMX_REQUIRE(out != nullptr, "filter_create_custom: out is NULL");
        *out = nullptr;
        amx::Filter *f = amx::filter_build_custom(n_det, n_rows, real_space, data, width);
        *out = new astra_amd_filter_s{f};
    });
}

int astra_amd_filter_create(unsigned int n_det, const float *half_spectrum, unsigned int n_bins,
                            astra_amd_filter *out) {
    return guarded([&] {
        AMX_REQUIRE(out != nullptr, "filter_create: out is NULL");
        *out = nullptr;
        amx::Filter *f = amx::filter_build(n_det, half_spectrum, n_bins);
        *out = new astra_amd_filter_s{f};
    });
}

int astra_amd_filter_destroy(astra_amd_filter f) {
    return guarded([&] {
        if (!f) return;
        delete f->f;
        delete f;
    });
}

int astra_amd_filter_rows(astra_amd_filter f, float *rows, size_t n_rows, void *stream) {
    return guarded([&] {
        AMX_REQUIRE(f && f->f, "filter handle is NULL");
        AMX_REQUIRE(rows != nullptr || n_rows == 0, "filter_rows: NULL data pointer");
        check_device(f->f->device);
        amx::filter_rows_launch(*f->f, rows, n_rows, nullptr, 1, as_stream(stream));
    });
}

int astra_amd_fdk_preweight(astra_amd_geom3d g, float *proj, void *stream) {
    return guarded([&] {
        AMX_REQUIRE(g && g->g, "geometry handle is NULL");
        AMX_REQUIRE(proj != nullptr, "fdk_preweight: NULL data pointer");
        check_device(g->g->device);
        amx::fdk_preweight_only(*g->g, proj, as_stream(stream));
    });
}

int astra_amd_fdk(astra_amd_geom3d g, float *proj, float *vol, float output_scale,
                  int short_scan, astra_amd_filter f, int mode, void *stream) {
    return guarded([&] {
        AMX_REQUIRE(g && g->g, "geometry handle is NULL");
        AMX_REQUIRE(proj && vol, "fdk: NULL data pointer");
        check_device(g->g->device);
        amx::fdk_run(*g->g, proj, vol, output_scale, f ? f->f : nullptr, short_scan != 0,
                     out_mode(mode), as_stream(stream));
    });
}

int astra_amd_fdk_filter(astra_amd_geom3d g, float *proj, int short_scan, astra_amd_filter f,
                         void *stream) {
    return guarded([&] {
        AMX_REQUIRE(g && g->g, "geometry handle is NULL");
        AMX_REQUIRE(proj, "fdk_filter: NULL data pointer");
        check_device(g->g->device);
        amx::fdk_filter_run(*g->g, proj, f ? f->f : nullptr, short_scan != 0, as_stream(stream));
    });
}

int astra_amd_fbp2d(astra_amd_geom2d g, float *sino, float *vol, float output_scale,
                    int short_scan, astra_amd_filter f, int mode, void *stream) {
    return guarded([&] {
        AMX_REQUIRE(g && g->g, "geometry handle is NULL");
        AMX_REQUIRE(sino && vol, "fbp2d: NULL data pointer");
        check_device(g->g->device);
        amx::fbp2d_run(*g->g, sino, vol, output_scale, f ? f->f : nullptr, short_scan != 0,
                       out_mode(mode), as_stream(stream));
    });
}

// ------------------------------------------------------ batched 2D -----
int astra_amd_batch2d_create(unsigned int n_slices, const int *kinds, const unsigned int *widths,
                             const unsigned int *heights, const unsigned int *n_angles,
                             const unsigned int *n_dets, const float *vecs,
                             const float *output_scales, astra_amd_batch2d *out) {
    return guarded([&] {
        AMX_REQUIRE(out != nullptr, "batch2d_create: out is NULL");
        *out = nullptr;
        amx::Batch2D *b = amx::batch2d_build(n_slices, kinds, widths, heights, n_angles, n_dets,
                                             vecs, output_scales);
        int dev = 0;
        (void)hipGetDevice(&dev);
        *out = new astra_amd_batch2d_s{b, dev};
    });
}

int astra_amd_batch2d_offsets(astra_amd_batch2d b, size_t *vol_offsets, size_t *sino_offsets) {
    return guarded([&] {
        AMX_REQUIRE(b && b->b, "batch handle is NULL");
        amx::batch2d_offsets(*b->b, vol_offsets, sino_offsets);
    });
}

int astra_amd_batch2d_fp(astra_amd_batch2d b, const float *vols, float *sinos, void *stream) {
    return guarded([&] {
        AMX_REQUIRE(b && b->b, "batch handle is NULL");
        check_device(b->device);
        amx::batch2d_fp(*b->b, vols, sinos, as_stream(stream));
    });
}

int astra_amd_batch2d_sirt(astra_amd_batch2d b, const float *sinos, float *vols, int iterations,
                           float relaxation, float min_constraint, float max_constraint,
                           int use_graph, void *stream) {
    return guarded([&] {
        AMX_REQUIRE(b && b->b, "batch handle is NULL");
        check_device(b->device);
        amx::batch2d_sirt(*b->b, sinos, vols, iterations, relaxation, min_constraint,
                          max_constraint, use_graph, as_stream(stream));
    });
}

int astra_amd_batch2d_destroy(astra_amd_batch2d b) {
    return guarded([&] {
        if (!b) return;
        amx::batch2d_destroy(b->b);
        delete b;
    });
}

// ---------------------------------------------------------- profiling ---
int astra_amd_profile_enable(int on) {
    return guarded([&] { amx::prof_enable(on != 0); });
}

int astra_amd_profile_read(const char *kernel, double *total_ms, double *units,
                           unsigned long long *launches) {
    return guarded([&] {
        AMX_REQUIRE(kernel && total_ms && units && launches, "profile_read: NULL argument");
        amx::prof_read(kernel, *total_ms, *units, *launches);
    });
}

int astra_amd_scratch_usage(size_t *in_use, size_t *peak, int reset_peak) {
    return guarded([&] {
        AMX_REQUIRE(in_use && peak, "scratch_usage: NULL argument");
        amx::scratch_usage(*in_use, *peak, reset_peak != 0);
    });
}

int astra_amd_calibrate_gather(int mode, unsigned int iterations, double *lane_loads_per_second,
                               void *stream) {
    return guarded([&] {
        AMX_REQUIRE(lane_loads_per_second != nullptr, "calibrate_gather: out is NULL");
        *lane_loads_per_second = amx::gather_calibrate(mode, iterations, as_stream(stream));
    });
}

// -------------------------------------------------------- elementwise ---
int astra_amd_dot(const float *a, const float *b, size_t n, double *out, void *stream) {
    return guarded([&] {
        AMX_REQUIRE(out != nullptr, "dot: out is NULL");
        AMX_REQUIRE((a && b) || n == 0, "dot: NULL data pointer");
        hipStream_t s = as_stream(stream);
        amx::Scratch r(sizeof(double), s);
        if (n == 0) {
            *out = 0.0;
            return;
        }
        amx::dot_launch(a, b, n, r.as<double>(), s);
        AMX_HIP(hipMemcpyAsync(out, r.as<double>(), sizeof(double), hipMemcpyDeviceToHost, s));
        AMX_HIP(hipStreamSynchronize(s));
    });
}

// -------------------------------------------------------- algorithms ----
int astra_amd_algo2d_create(int type, astra_amd_geom2d g, const astra_amd_params2d *p,
                            astra_amd_algo *out) {
    return guarded([&] {
        AMX_REQUIRE(out != nullptr, "algo2d_create: out is NULL");
        *out = nullptr;
        check2d(g, p);
        amx::AlgoBase *a = amx::algo_build(type, g->g, nullptr, p->output_scale, 0,
                                           ss_of(p->rays_per_det), ss_of(p->rays_per_pixel_dim));
        *out = new astra_amd_algo_s{a, g->g->device};
    });
}

int astra_amd_algo3d_create(int type, astra_amd_geom3d g, const astra_amd_params3d *p,
                            astra_amd_algo *out) {
    return guarded([&] {
        AMX_REQUIRE(out != nullptr, "algo3d_create: out is NULL");
        *out = nullptr;
        check3d(g, p);
        AMX_REQUIRE(p->kernel == ASTRA_AMD_KER3D_DEFAULT || p->kernel == ASTRA_AMD_KER3D_2D_WEIGHTING,
                    "algo3d: projection kernel %d not usable for SIRT/CGLS", p->kernel);
        if (g->g->kind == ASTRA_AMD_GEOM_CYLCONE) {
            // src/CudaSirtAlgorithm3D.cpp:155-159, src/CudaCglsAlgorithm3D.cpp:156-160
            AMX_REQUIRE(ss_of(p->rays_per_det_dim) == 1 && ss_of(p->rays_per_voxel_dim) == 1,
                        "Detector/voxel supersampling is not supported for cyl_cone_vec geometry.");
            AMX_REQUIRE(p->kernel == ASTRA_AMD_KER3D_DEFAULT,
                        "algo3d: the cylindrical-detector geometry supports the default kernel only");
        }
        amx::AlgoBase *a = amx::algo_build(type, nullptr, g->g, p->output_scale, p->kernel,
                                           ss_of(p->rays_per_det_dim), ss_of(p->rays_per_voxel_dim));
        *out = new astra_amd_algo_s{a, g->g->device};
    });
}

int astra_amd_algo_set_option(astra_amd_algo a, const char *key, double value) {
    return guarded([&] {
        AMX_REQUIRE(a && a->a, "algorithm handle is NULL");
        amx::algo_set_option(a->a, key, value);
    });
}

int astra_amd_algo_set_order(astra_amd_algo a, const int *order, unsigned int n) {
    return guarded([&] {
        AMX_REQUIRE(a && a->a, "algorithm handle is NULL");
        amx::algo_set_order(a->a, order, n);
    });
}

int astra_amd_algo_set_mask(astra_amd_algo a, const char *key, const float *mask) {
    return guarded([&] {
        AMX_REQUIRE(a && a->a, "algorithm handle is NULL");
        amx::algo_set_mask(a->a, key, mask);
    });
}

int astra_amd_algo_run(astra_amd_algo a, const float *sino, float *vol, int iterations,
                       void *stream) {
    return guarded([&] {
        AMX_REQUIRE(a && a->a, "algorithm handle is NULL");
        AMX_REQUIRE(sino && vol, "algo_run: NULL data pointer");
        AMX_REQUIRE(iterations >= 0, "algo_run: negative iteration count");
        check_device(a->device);
        amx::algo_run(a->a, sino, vol, iterations, as_stream(stream));
    });
}

int astra_amd_set_abort_hook(astra_amd_abort_hook fn, void *user) {
    return guarded([&] { amx::set_abort_hook(fn, user); });
}

int astra_amd_algo_residual_norm(astra_amd_algo a, const float *sino, const float *vol,
                                 double *out, void *stream) {
    return guarded([&] {
        AMX_REQUIRE(a && a->a && out, "algo_residual_norm: NULL argument");
        AMX_REQUIRE(sino && vol, "algo_residual_norm: NULL data pointer");
        check_device(a->device);
        *out = amx::algo_residual(a->a, sino, vol, as_stream(stream));
    });
}

int astra_amd_algo_destroy(astra_amd_algo a) {
    return guarded([&] {
        if (!a) return;
        amx::algo_destroy(a->a);
        delete a;
    });
}

// ------------------------------------------------- pitched buffers -----
int astra_amd_fp2d_pitched(astra_amd_geom2d g, const float *vol, size_t vol_pitch, float *sino,
                           size_t sino_pitch, const astra_amd_params2d *p, int mode,
                           void *stream) {
    return guarded([&] {
        check2d(g, p);
        AMX_REQUIRE(vol && sino, "fp2d: NULL data pointer");
        const amx::Geom2D &G = *g->g;
        hipStream_t st = as_stream(stream);
        Contig v(vol, G.w, G.h, vol_pitch, true, st);
        Contig o(sino, G.nd, G.na, sino_pitch, mode == ASTRA_AMD_MODE_ADD, st);
        int rc = astra_amd_fp2d(g, v.get(), o.get(), p, mode, stream);
        if (rc) amx::fail(rc, "%s", t_err.c_str());
        o.store();
    });
}

int astra_amd_bp2d_pitched(astra_amd_geom2d g, const float *sino, size_t sino_pitch, float *vol,
                           size_t vol_pitch, const astra_amd_params2d *p, int mode,
                           void *stream) {
    return guarded([&] {
        check2d(g, p);
        AMX_REQUIRE(vol && sino, "bp2d: NULL data pointer");
        const amx::Geom2D &G = *g->g;
        hipStream_t st = as_stream(stream);
        Contig i(sino, G.nd, G.na, sino_pitch, true, st);
        Contig o(vol, G.w, G.h, vol_pitch, mode == ASTRA_AMD_MODE_ADD, st);
        int rc = astra_amd_bp2d(g, i.get(), o.get(), p, mode, stream);
        if (rc) amx::fail(rc, "%s", t_err.c_str());
        o.store();
    });
}

int astra_amd_fp3d_pitched(astra_amd_geom3d g, const float *vol, size_t vol_pitch, float *proj,
                           size_t proj_pitch, const astra_amd_params3d *p, int mode,
                           void *stream) {
    return guarded([&] {
        check3d(g, p);
        AMX_REQUIRE(vol && proj, "fp3d: NULL data pointer");
        const amx::Geom3D &G = *g->g;
        hipStream_t st = as_stream(stream);
        Contig v(vol, G.nx, (size_t)G.ny * G.nz, vol_pitch, true, st);
        Contig o(proj, G.nu, (size_t)G.na * G.nv, proj_pitch, mode == ASTRA_AMD_MODE_ADD, st);
        int rc = astra_amd_fp3d(g, v.get(), o.get(), p, mode, stream);
        if (rc) amx::fail(rc, "%s", t_err.c_str());
        o.store();
    });
}

int astra_amd_bp3d_pitched(astra_amd_geom3d g, const float *proj, size_t proj_pitch, float *vol,
                           size_t vol_pitch, const astra_amd_params3d *p, int mode,
                           void *stream) {
    return guarded([&] {
        check3d(g, p);
        AMX_REQUIRE(vol && proj, "bp3d: NULL data pointer");
        const amx::Geom3D &G = *g->g;
        hipStream_t st = as_stream(stream);
        Contig i(proj, G.nu, (size_t)G.na * G.nv, proj_pitch, true, st);
        Contig o(vol, G.nx, (size_t)G.ny * G.nz, vol_pitch, mode == ASTRA_AMD_MODE_ADD, st);
        int rc = astra_amd_bp3d(g, i.get(), o.get(), p, mode, stream);
        if (rc) amx::fail(rc, "%s", t_err.c_str());
        o.store();
    });
}

}  // extern "C"
