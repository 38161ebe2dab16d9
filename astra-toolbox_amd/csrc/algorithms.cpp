// Reconstruction algorithm objects over device data (SIRT, CGLS in 2D and
// 3D) and the FDK / FBP drivers.
//
// The iteration state lives in projector-friendly layouts: residuals are
// kept in the zero-padded angle-major layout that the back projector reads,
// the forward projector writes them there directly with the residual and
// line weights fused into its epilogue, and the back projector applies the
// SIRT update (pixel weights, relaxation, clamping) in its epilogue.  One
// SIRT iteration is therefore two projector launches plus the forward
// projector's input copy; no host synchronisation happens inside run().
#include <atomic>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <string>

#include "devutil.hpp"
#include "internal.hpp"

namespace amx {

// ------------------------------------------------------------ operator ---
// Uniform view of a 2D or 3D geometry for the algorithm loops.
struct Op {
    Geom2D *g2 = nullptr;
    Geom3D *g3 = nullptr;
    float out_scale = 1.f;
    int kernel = ASTRA_AMD_KER3D_DEFAULT;
    unsigned rays = 1;  // FP: rays per detector (2D) / per detector dimension (3D)
    unsigned vss = 1;   // BP: sub-pixels / sub-voxels per dimension

    size_t nvol() const {
        return g3 ? (size_t)g3->nx * g3->ny * g3->nz : (size_t)g2->w * g2->h;
    }
    size_t nproj() const {
        return g3 ? (size_t)g3->nv * g3->na * g3->nu : (size_t)g2->na * g2->nd;
    }
    size_t npad() const { return g3 ? padded_am_elems(*g3) : padded_sino_elems(*g2); }

    // out_padded = mode(scale * A x) with the residual epilogue when mode is OUT_RESID
    void fp(const float *x, float *out_padded, float scale, int mode, const float *b,
            const float *w, hipStream_t s) const {
        if (g3)
            fp3d_launch(*g3, x, out_padded, out_scale * scale, mode, b, w, true, s, rays);
        else
            fp2d_launch(*g2, x, out_padded, out_scale * scale, mode, b, w, true, rays, s);
    }
    void bp(const float *pp, float *x, float scale, int mode, const float *uw, float lo,
            float hi, hipStream_t s) const {
        if (g3) {
            int k = (kernel == ASTRA_AMD_KER3D_2D_WEIGHTING) ? kernel : ASTRA_AMD_KER3D_DEFAULT;
            bp3d_launch_padded(*g3, pp, x, out_scale * scale, k, mode, uw, lo, hi, s, vss);
        } else {
            bp2d_launch_padded(*g2, pp, x, out_scale * scale, false, mode, uw, lo, hi, s, vss);
        }
    }
    void to_padded(const float *proj, float *pp, hipStream_t s) const {
        if (g3)
            proj_to_padded_am(*g3, proj, pp, s);
        else
            sino_to_padded(*g2, proj, pp, s);
    }
};

// The abort hook (setShouldAbortHook) and the per-run gate that asks it:
// before iteration it, wait for iteration it - kDepth to finish on the GPU
// (so a stop request lands within kDepth iterations of GPU work), then ask.
std::atomic<astra_amd_abort_hook> g_abort_hook{nullptr};
std::atomic<void *> g_abort_user{nullptr};

struct AbortGate {
    static constexpr int kDepth = 3;
    hipStream_t s;
    astra_amd_abort_hook fn;
    void *user;
    hipEvent_t ev[kDepth] = {};
    explicit AbortGate(hipStream_t s_) : s(s_), fn(g_abort_hook.load()), user(g_abort_user.load()) {
        if (fn)
            for (auto &e : ev) AMX_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    ~AbortGate() {
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
    }
    AbortGate(const AbortGate &) = delete;
    AbortGate &operator=(const AbortGate &) = delete;
    // true: stop before iteration it
    bool stop(int it) {
        if (!fn) return false;
        if (it >= kDepth) AMX_HIP(hipEventSynchronize(ev[it % kDepth]));
        return fn(user) != 0;
    }
    void done(int it) {
        if (fn) AMX_HIP(hipEventRecord(ev[it % kDepth], s));
    }
    [[noreturn]] void aborted(int it) const {
        fail(ASTRA_AMD_ERR_ABORTED, "run aborted by the abort hook after %d iterations", it);
    }
};

struct AlgoBase {
    Op op;
    int type = ASTRA_AMD_ALGO_SIRT;
    double relaxation = 1.0;
    float lo = -std::numeric_limits<float>::infinity();
    float hi = std::numeric_limits<float>::infinity();
    const float *recon_mask = nullptr, *sino_mask = nullptr;
    const float *min_mask = nullptr, *max_mask = nullptr;
    bool weights_ready = false;
    int device = 0;
    DevArray line_w, pixel_w, r, w, p, z, t, scal;
    std::mutex mu;

    virtual ~AlgoBase() = default;
    virtual void run(const float *b, float *x, int iters, hipStream_t s) = 0;

    double residual(const float *b, const float *x, hipStream_t s) {
        ensure_buffers();
        const float *xin = x;
        if (recon_mask) {
            copy_launch(t.as<float>(), x, op.nvol(), s);
            mul_launch(t.as<float>(), recon_mask, op.nvol(), s);
            xin = t.as<float>();
        }
        op.fp(xin, r.as<float>(), -1.f, OUT_RESID, b, nullptr, s);
        dot_launch(r.as<float>(), r.as<float>(), op.npad(), scal.as<double>(), s);
        double h = 0;
        AMX_HIP(hipMemcpyAsync(&h, scal.as<double>(), sizeof(double), hipMemcpyDeviceToHost, s));
        AMX_HIP(hipStreamSynchronize(s));
        return std::sqrt(std::max(h, 0.0));
    }

    void ensure_buffers() {
        if (r.bytes()) return;
        const size_t nv = op.nvol() * sizeof(float), np = op.npad() * sizeof(float);
        r = DevArray(np);
        t = DevArray(std::max(nv, op.nproj() * sizeof(float)));
        scal = DevArray(4 * sizeof(double));
        // padded buffers must have zero borders: they are only ever written
        // in their interior
        AMX_HIP(hipMemset(r.as<float>(), 0, np));
        if (type != ASTRA_AMD_ALGO_CGLS) {
            line_w = DevArray(np);
            pixel_w = DevArray(nv);
            AMX_HIP(hipMemset(line_w.as<float>(), 0, np));
        } else {
            w = DevArray(np);
            p = DevArray(nv);
            z = DevArray(nv);
            AMX_HIP(hipMemset(w.as<float>(), 0, np));
        }
    }
};

// SIRT: CCudaSirtAlgorithm::precomputeWeights / run
// (src/CudaSirtAlgorithm.cpp:187-222, 347-379) and
// CCudaSirtAlgorithm3D (src/CudaSirtAlgorithm3D.cpp:261-376).
struct Sirt : AlgoBase {
    void precompute(hipStream_t s) {
        const size_t nv = op.nvol(), np = op.nproj();
        float *tv = t.as<float>();
        // line weights: inv(FP(mask or 1)) [x sinogram mask]
        if (recon_mask)
            copy_launch(tv, recon_mask, nv, s);
        else
            fill_launch(tv, nv, 1.f, s);
        op.fp(tv, line_w.as<float>(), 1.f, OUT_SET, nullptr, nullptr, s);
        invert_launch(line_w.as<float>(), op.npad(), 1.f, s);
        // pixel weights: lambda * inv(BP(sinogram mask or 1)) [x volume mask]
        if (sino_mask)
            copy_launch(tv, sino_mask, np, s);
        else
            fill_launch(tv, np, 1.f, s);
        op.to_padded(tv, r.as<float>(), s);
        if (sino_mask) {
            // line weights x sinogram mask (padded layout)
            mul_launch(line_w.as<float>(), r.as<float>(), op.npad(), s);
        }
        op.bp(r.as<float>(), pixel_w.as<float>(), 1.f, OUT_SET, nullptr, lo, hi, s);
        invert_launch(pixel_w.as<float>(), nv, (float)relaxation, s);
        if (recon_mask) mul_launch(pixel_w.as<float>(), recon_mask, nv, s);
    }

    void run(const float *b, float *x, int iters, hipStream_t s) override {
        ensure_buffers();
        if (!weights_ready || recon_mask || sino_mask) {
            precompute(s);
            weights_ready = true;
        }
        const size_t nv = op.nvol();
        AbortGate gate(s);
        for (int it = 0; it < iters; ++it) {
            if (gate.stop(it)) gate.aborted(it);
            const float *xin = x;
            if (recon_mask) {
                copy_launch(t.as<float>(), x, nv, s);
                mul_launch(t.as<float>(), recon_mask, nv, s);
                xin = t.as<float>();
            }
            // r = (b - A x) * lineW, written in the padded BP layout
            op.fp(xin, r.as<float>(), -1.f, OUT_RESID, b, line_w.as<float>(), s);
            // x = clamp(x + pixelW * A^T r)
            op.bp(r.as<float>(), x, 1.f, OUT_UPDATE, pixel_w.as<float>(), lo, hi, s);
            if (min_mask || max_mask)
                clamp_launch(x, nv, -INFINITY, INFINITY, min_mask, max_mask, s);
            gate.done(it);
        }
    }
};

// CGLS: CCudaCglsAlgorithm::run (src/CudaCglsAlgorithm.cpp:166-266) and
// CCudaCglsAlgorithm3D (src/CudaCglsAlgorithm3D.cpp:264-360); restarts on
// every run().  The scalars gamma, <w,w> stay in device memory (float64).
struct Cgls : AlgoBase {
    void run(const float *b, float *x, int iters, hipStream_t s) override {
        ensure_buffers();
        const size_t nv = op.nvol(), np = op.npad();
        double *g0 = scal.as<double>(), *g1 = g0 + 1, *ww = g0 + 2;
        const float *xin = x;
        if (recon_mask) {
            copy_launch(t.as<float>(), x, nv, s);
            mul_launch(t.as<float>(), recon_mask, nv, s);
            xin = t.as<float>();
        }
        // r = b - A x ; p = A^T r ; gamma = <p, p>
        op.fp(xin, r.as<float>(), -1.f, OUT_RESID, b, nullptr, s);
        op.bp(r.as<float>(), p.as<float>(), 1.f, OUT_SET, nullptr, lo, hi, s);
        if (recon_mask) mul_launch(p.as<float>(), recon_mask, nv, s);
        dot_launch(p.as<float>(), p.as<float>(), nv, g0, s);
        AbortGate gate(s);
        for (int it = 0; it < iters; ++it) {
            if (gate.stop(it)) gate.aborted(it);
            // w = A p ; alpha = gamma / <w, w>
            op.fp(p.as<float>(), w.as<float>(), 1.f, OUT_SET, nullptr, nullptr, s);
            dot_launch(w.as<float>(), w.as<float>(), np, ww, s);
            axpy_ratio_launch(x, p.as<float>(), nv, g0, ww, 1.0, s);
            axpy_ratio_launch(r.as<float>(), w.as<float>(), np, g0, ww, -1.0, s);
            // z = A^T r ; beta = <z,z> / gamma ; p = z + beta p
            op.bp(r.as<float>(), z.as<float>(), 1.f, OUT_SET, nullptr, lo, hi, s);
            if (recon_mask) mul_launch(z.as<float>(), recon_mask, nv, s);
            dot_launch(z.as<float>(), z.as<float>(), nv, g1, s);
            xpby_ratio_launch(p.as<float>(), z.as<float>(), nv, g1, g0, s);
            std::swap(g0, g1);
            gate.done(it);
        }
    }
};

// EM: CCudaEMAlgorithm::precomputeWeights / run (src/CudaEMAlgorithm.cpp:
// 149-253): pixelW = inv(A^T 1); x *= pixelW A^T (b / A x), the ratio set
// to 0 where A x <= 1e-6.  Both steps fuse into the projector epilogues.
struct Em : AlgoBase {
    void run(const float *b, float *x, int iters, hipStream_t s) override {
        AMX_REQUIRE(!recon_mask && !sino_mask && !min_mask && !max_mask,
                    "EM: mask options are not supported");
        ensure_buffers();
        if (!weights_ready) {
            const size_t np = op.nproj();
            fill_launch(t.as<float>(), np, 1.f, s);
            op.to_padded(t.as<float>(), r.as<float>(), s);
            op.bp(r.as<float>(), pixel_w.as<float>(), 1.f, OUT_SET, nullptr, lo, hi, s);
            invert_launch(pixel_w.as<float>(), op.nvol(), 1.f, s);
            weights_ready = true;
        }
        AbortGate gate(s);
        for (int it = 0; it < iters; ++it) {
            if (gate.stop(it)) gate.aborted(it);
            op.fp(x, r.as<float>(), 1.f, OUT_RATIO, b, nullptr, s);
            op.bp(r.as<float>(), x, 1.f, OUT_EMUPD, pixel_w.as<float>(), lo, hi, s);
            gate.done(it);
        }
    }
};

// SART: CCudaSartAlgorithm (src/CudaSartAlgorithm.cpp:196-316), 2D only.
// One angle per iteration, in the given projection order (or 0, 1, 2, ...),
// continuing across run() calls: r_a = (b_a - A_a x) lineW_a;
// x = clamp(x + lambda * mask * A_a^T r_a) with SART's unweighted BP.
struct Sart : AlgoBase {
    std::vector<int> order;
    long long iteration = 0;

    void precompute(hipStream_t s) {
        const size_t nv = op.nvol();
        float *tv = t.as<float>();
        if (recon_mask)
            copy_launch(tv, recon_mask, nv, s);
        else
            fill_launch(tv, nv, 1.f, s);
        op.fp(tv, line_w.as<float>(), 1.f, OUT_SET, nullptr, nullptr, s);
        invert_launch(line_w.as<float>(), op.npad(), 1.f, s);
    }

    void run(const float *b, float *x, int iters, hipStream_t s) override {
        AMX_REQUIRE(op.g2 != nullptr, "SART: 2D geometries only");
        AMX_REQUIRE(!sino_mask, "SART: the sinogram mask option is not supported");
        ensure_buffers();
        if (!weights_ready || recon_mask) {
            precompute(s);
            weights_ready = true;
        }
        const Geom2D &g = *op.g2;
        const size_t nv = op.nvol();
        const size_t pitch = g.nd + 2 * 2;
        AbortGate gate(s);
        for (int it = 0; it < iters; ++it, ++iteration) {
            if (gate.stop(it)) gate.aborted(it);
            const int a = order.empty() ? (int)(iteration % g.na)
                                        : order[(size_t)(iteration % (long long)order.size())];
            const float *xin = x;
            if (recon_mask) {
                copy_launch(t.as<float>(), x, nv, s);
                mul_launch(t.as<float>(), recon_mask, nv, s);
                xin = t.as<float>();
            }
            fp2d_launch_range(g, xin, r.as<float>(), -op.out_scale, OUT_RESID, b,
                              line_w.as<float>(), true, op.rays, (unsigned)a, 1, s);
            bp2d_sart_launch(g, r.as<float>() + (size_t)a * pitch, x, (unsigned)a,
                             op.out_scale * (float)relaxation, recon_mask, lo, hi, s);
            if (min_mask || max_mask)
                clamp_launch(x, nv, -INFINITY, INFINITY, min_mask, max_mask, s);
            gate.done(it);
        }
    }
};

void algo_set_order(AlgoBase *a, const int *order, unsigned n) {
    std::lock_guard<std::mutex> lk(a->mu);
    auto *sart = dynamic_cast<Sart *>(a);
    AMX_REQUIRE(sart != nullptr, "ProjectionOrder applies to SART only");
    AMX_REQUIRE(order != nullptr || n == 0, "ProjectionOrder: NULL list");
    const unsigned na = a->op.g2->na;
    for (unsigned i = 0; i < n; ++i)
        AMX_REQUIRE(order[i] >= 0 && (unsigned)order[i] < na,
                    "ProjectionOrderList entry %d out of range [0, %u)", order[i], na);
    sart->order.assign(order, order + n);
}

AlgoBase *algo_build(int type, Geom2D *g2, Geom3D *g3, float out_scale, int kernel,
                     unsigned rays, unsigned vss) {
    AlgoBase *a = nullptr;
    if (type == ASTRA_AMD_ALGO_SIRT)
        a = new Sirt();
    else if (type == ASTRA_AMD_ALGO_CGLS)
        a = new Cgls();
    else if (type == ASTRA_AMD_ALGO_EM)
        a = new Em();
    else if (type == ASTRA_AMD_ALGO_SART) {
        AMX_REQUIRE(g2 != nullptr, "SART: 2D geometries only");
        AMX_REQUIRE(vss == 1, "SART: VoxelSuperSampling is not supported");
        a = new Sart();
    } else
        fail(ASTRA_AMD_ERR_ARG, "algorithm: unknown type %d", type);
    a->type = type;
    a->op.g2 = g2;
    a->op.g3 = g3;
    a->op.out_scale = out_scale;
    a->op.kernel = kernel;
    a->op.rays = rays;
    a->op.vss = vss;
    AMX_HIP(hipGetDevice(&a->device));
    return a;
}

void algo_destroy(AlgoBase *a) { delete a; }

void algo_set_option(AlgoBase *a, const char *key, double v) {
    std::string k(key ? key : "");
    if (k == "Relaxation") {
        a->relaxation = v;
        a->weights_ready = false;
    } else if (k == "MinConstraint") {
        a->lo = (float)v;
    } else if (k == "MaxConstraint") {
        a->hi = (float)v;
    } else {
        fail(ASTRA_AMD_ERR_ARG, "algorithm: unknown option '%s'", k.c_str());
    }
}

void algo_set_mask(AlgoBase *a, const char *key, const float *m) {
    std::string k(key ? key : "");
    if (k == "ReconstructionMask") a->recon_mask = m;
    else if (k == "SinogramMask") a->sino_mask = m;
    else if (k == "MinMask") a->min_mask = m;
    else if (k == "MaxMask") a->max_mask = m;
    else fail(ASTRA_AMD_ERR_ARG, "algorithm: unknown mask '%s'", k.c_str());
    a->weights_ready = false;
}

void set_abort_hook(astra_amd_abort_hook fn, void *user) {
    g_abort_user.store(user);
    g_abort_hook.store(fn);
}

void algo_run(AlgoBase *a, const float *b, float *x, int iters, hipStream_t s) {
    std::lock_guard<std::mutex> lk(a->mu);
    a->run(b, x, iters, s);
}

double algo_residual(AlgoBase *a, const float *b, const float *x, hipStream_t s) {
    std::lock_guard<std::mutex> lk(a->mu);
    return a->residual(b, x, s);
}

// ----------------------------------------------------------------- FDK ---
namespace {
void upload_twiddles(Filter &f) {
    const unsigned n = f.n_fft;
    std::vector<float> tw(2 * (size_t)n);  // n complex: exp(-2 pi i k / n), k < n
    for (unsigned k = 0; k < n; ++k) {
        double ang = -2.0 * M_PI * (double)k / (double)n;
        tw[2 * k] = (float)std::cos(ang);
        tw[2 * k + 1] = (float)std::sin(ang);
    }
    f.twiddle = DevArray(sizeof(float) * tw.size());
    f.twiddle.upload(tw.data(), sizeof(float) * tw.size());
}
}  // namespace

Filter *filter_build(unsigned n_det, const float *half, unsigned n_bins) {
    AMX_REQUIRE(n_det > 0, "filter: n_det must be positive");
    unsigned n = next_pow2(2 * n_det);
    std::unique_ptr<Filter> f(new Filter());
    f->n_det = n_det;
    f->n_fft = n;
    AMX_HIP(hipGetDevice(&f->device));
    if (half == nullptr && n_bins == 0) {  // identity ('none')
        f->identity = true;
        return f.release();
    }
    AMX_REQUIRE(n_bins == n / 2 + 1, "filter: expected %u spectrum bins for %u detectors, got %u",
                n / 2 + 1, n_det, n_bins);
    AMX_REQUIRE(half != nullptr, "filter: spectrum is NULL");
    std::vector<float> full(n);
    for (unsigned k = 0; k < n; ++k) full[k] = half[k <= n / 2 ? k : n - k];
    f->spec = DevArray(sizeof(float) * n);
    f->spec.upload(full.data(), sizeof(float) * n);
    upload_twiddles(*f);
    return f.release();
}

Filter *filter_build_custom(unsigned n_det, unsigned n_rows, int real_space, const float *data,
                            unsigned width) {
    AMX_REQUIRE(n_det > 0 && n_rows > 0, "filter: n_det and n_rows must be positive");
    AMX_REQUIRE(data != nullptr, "filter: data is NULL");
    std::vector<float> spec;
    custom_filter_spectra(n_det, n_rows, real_space, data, width, spec);
    std::unique_ptr<Filter> f(new Filter());
    f->n_det = n_det;
    f->n_fft = next_pow2(2 * n_det);
    f->general = true;
    f->n_spec = n_rows;
    AMX_HIP(hipGetDevice(&f->device));
    f->spec = DevArray(sizeof(float) * spec.size());
    f->spec.upload(spec.data(), sizeof(float) * spec.size());
    upload_twiddles(*f);
    return f.release();
}

Filter *ramlak_filter(unsigned n_det) {
    std::vector<double> h;
    ramlak_half_spectrum(n_det, 1.0, h);
    std::vector<float> hf(h.begin(), h.end());
    return filter_build(n_det, hf.data(), (unsigned)hf.size());
}

// FDK pre-weights and default filter of a geometry, built once
static void fdk_prepare(const Geom3D &g) {
    std::lock_guard<std::mutex> lk(g.fdk_mu);
    if (g.fdk_filter) return;
    std::vector<float> w;
    fdk_weights(g, w);
    DevArray dw(sizeof(float) * w.size());
    dw.upload(w.data(), sizeof(float) * w.size());
    g.fdk_w = std::move(dw);
    g.fdk_filter.reset(ramlak_filter(g.nu));
}

// astraCUDA3d::FDK (cuda/3d/fdk.cu:415-478): pre-weight, filter, FDK BP
void fdk_filter_run(const Geom3D &g, float *proj, const Filter *f, bool short_scan,
                    hipStream_t s) {
    AMX_REQUIRE(g.kind == ASTRA_AMD_GEOM_CONE, "FDK requires a cone geometry");
    fdk_prepare(g);
    const Filter *flt = f ? f : g.fdk_filter.get();
    AMX_REQUIRE(flt->n_det == g.nu, "FDK: filter built for %u detectors, geometry has %u",
                flt->n_det, g.nu);
    AMX_REQUIRE(flt->n_spec == 1 || flt->n_spec == g.na,
                "FDK: a per-angle filter needs %u rows, got %u", g.na, flt->n_spec);
    const float *parker = nullptr;
    if (short_scan && g.na > 1) {
        std::lock_guard<std::mutex> lk(g.fdk_mu);
        if (!g.fdk_parker.bytes()) {
            std::vector<float> w;
            fdk_parker(g, w);
            DevArray dp(sizeof(float) * w.size());
            dp.upload(w.data(), sizeof(float) * w.size());
            g.fdk_parker = std::move(dp);
        }
        parker = g.fdk_parker.as<float>();
    }
    // fused cosine (and Parker) pre-weights + ramp filter over all nv * na rows
    filter_rows_launch(*flt, proj, (size_t)g.nv * g.na, g.fdk_w.as<float>(), g.na, s, parker,
                       g.na);
}

void fdk_run(const Geom3D &g, float *proj, float *vol, float out_scale, const Filter *f,
             bool short_scan, int mode, hipStream_t s) {
    fdk_filter_run(g, proj, f, short_scan, s);
    bp3d_launch(g, proj, vol, out_scale, ASTRA_AMD_KER3D_FDK_WEIGHTING, mode, nullptr,
                -INFINITY, INFINITY, s);
}

// FDK_PreWeight alone (cuda/3d/fdk.cu:168-264, full scan)
void fdk_preweight_only(const Geom3D &g, float *proj, hipStream_t s) {
    AMX_REQUIRE(g.kind == ASTRA_AMD_GEOM_CONE, "FDK pre-weighting requires a cone geometry");
    fdk_prepare(g);
    rowscale_launch(proj, (size_t)g.nv * g.na, g.nu, g.fdk_w.as<float>(), g.na, s);
}

// ----------------------------------------------------------------- FBP ---
// CCudaFilteredBackProjectionAlgorithm / astraCUDA::FBP
// (src/CudaFilteredBackProjectionAlgorithm.cpp:174-239, cuda/2d/fbp.cu:77-190)
void fbp2d_run(const Geom2D &g, float *sino, float *vol, float out_scale, const Filter *f,
               bool short_scan, int mode, hipStream_t s) {
    if (!f) {
        std::lock_guard<std::mutex> lk(g.fbp_mu);
        if (!g.fbp_ramlak) g.fbp_ramlak.reset(ramlak_filter(g.nd));
        f = g.fbp_ramlak.get();
    }
    AMX_REQUIRE(f->n_det == g.nd, "FBP: filter built for %u detectors, geometry has %u",
                f->n_det, g.nd);
    AMX_REQUIRE(f->n_spec == 1 || f->n_spec == g.na,
                "FBP: a per-angle filter needs %u rows, got %u", g.na, f->n_spec);
    if (g.kind == ASTRA_AMD_GEOM_FAN) {
        const float *parker = nullptr;
        {
            std::lock_guard<std::mutex> lk(g.fbp_mu);
            // one-row cone pre-weighting with getFanParameters of the last
            // angle (cuda/2d/fbp.cu:112-148, src/GeometryUtil2D.cpp:138-177)
            const float *p = &g.vecs[6 * (size_t)(g.na - 1)];
            double sx = p[0], sy = p[1], dsx = p[2], dsy = p[3], ux = p[4], uy = p[5];
            double sod = std::sqrt(sx * sx + sy * sy);
            double du = std::sqrt(ux * ux + uy * uy);
            double t = (-dsx * ux - dsy * uy) / (ux * ux + uy * uy);
            double odd = std::sqrt((dsx + t * ux) * (dsx + t * ux) + (dsy + t * uy) * (dsy + t * uy));
            if (!g.fbp_w.bytes()) {
                double D = sod + odd;
                double W = D * (D / (du * sod)) * (M_PI / 2.0) / (double)g.na;
                std::vector<float> w(g.nd);
                for (unsigned i = 0; i < g.nd; ++i) {
                    double U = (i - 0.5 * g.nd + 0.5) * du;
                    w[i] = (float)(W / std::sqrt(D * D + U * U));
                }
                DevArray dw(sizeof(float) * g.nd);
                dw.upload(w.data(), sizeof(float) * g.nd);
                g.fbp_w = std::move(dw);
            }
            if (short_scan && g.na > 1) {
                if (!g.fbp_parker.bytes()) {
                    // Parker weights with each angle's getFanParameters angle
                    // atan2(uy, ux) and the last angle's distances (cuda/2d/fbp.cu:117-144)
                    std::vector<double> ang(g.na);
                    for (unsigned a = 0; a < g.na; ++a)
                        ang[a] = std::atan2((double)g.vecs[6 * a + 5], (double)g.vecs[6 * a + 4]);
                    std::vector<float> pw;
                    parker_weights(ang, g.nd, sod, odd, du, pw);
                    DevArray dp(sizeof(float) * pw.size());
                    dp.upload(pw.data(), sizeof(float) * pw.size());
                    g.fbp_parker = std::move(dp);
                }
                parker = g.fbp_parker.as<float>();
            }
        }
        filter_rows_launch(*f, sino, g.na, g.fbp_w.as<float>(), g.na, s, parker, g.na);
        bp2d_launch(g, sino, vol, out_scale, true, mode, nullptr, -INFINITY, INFINITY, s);
    } else {
        filter_rows_launch(*f, sino, g.na, nullptr, 1, s);
        bp2d_launch(g, sino, vol, out_scale * (float)((M_PI / 2.0) / (double)g.na), false, mode,
                    nullptr, -INFINITY, INFINITY, s);
    }
}

}  // namespace amx
