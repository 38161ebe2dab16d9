// Reference-side binding of the 3D seam: new bodies for astraCUDA3d::FP,
// astraCUDA3d::BP and astraCUDA3d::FDK (include/astra/cuda/3d/mem3d.h:107-113)
// that call libastra_amd through its C ABI (include/astra_amd.h).
//
// This file belongs to an ASTRA source tree, not to libastra_amd: a
// maintainer builds it in place of cuda/3d/mem3d.cu's FP / BP / FDK (the
// reference's HIP configuration, ASTRA_BUILDING_HIP, so that the cuda*
// names of the headers map to HIP) and links -lastra_amd.  The product
// never includes or builds it; a CPU test compiles it with -fsyntax-only
// against the reference headers.
//
// Replaced bodies: cuda/3d/mem3d.cu:334-398 (FP), :416-485 (BP),
// :486-516 (FDK).

#include "astra/cuda/gpu_runtime_wrapper.h"

#include "astra/cuda/3d/mem3d.h"
#include "astra/cuda/3d/mem3d_internal.h"
#include "astra/Data3D.h"
#include "astra/Filters.h"
#include "astra/GeometryUtil3D.h"
#include "astra/Logging.h"

#include <cstring>
#include <vector>

#include "astra_amd.h"

namespace {

// Pitched device storage of a data object.  CDataGPU::getPtr() returns the
// cudaPitchedPtr by value (mem3d_internal.h:58); storage that wraps a
// cudaArray (getArray() != nullptr, ptr.ptr == nullptr) is refused, as the
// reference's own FP / BP assert (cuda/3d/mem3d.cu:342-343).
bool pitched(const astra::CData3D *d, cudaPitchedPtr &out) {
    const astraCUDA::CDataGPU *s = dynamic_cast<const astraCUDA::CDataGPU *>(d->getStorage());
    if (!s) {
        ASTRA_ERROR("astra_amd seam: data object is not stored on the GPU");
        return false;
    }
    if (s->getArray() != nullptr) {
        ASTRA_ERROR("astra_amd seam: cudaArray-backed GPU storage is not supported");
        return false;
    }
    out = s->getPtr();
    return out.ptr != nullptr;
}

// 12 doubles per angle [s, dS, u, v] (13 with the radius for the
// cylindrical detector), already normalised by convertAstraGeometry.
bool vectors(const astra::Geometry3DParameters &g, int &kind, std::vector<double> &v) {
    const unsigned n = g.getDims().iProjAngles;
    if (const astra::SConeProjection *p = g.getCone()) {
        kind = ASTRA_AMD_GEOM_CONE;
        v.resize(12 * (size_t)n);
        for (unsigned i = 0; i < n; ++i) {
            const astra::SConeProjection &q = p[i];
            const double a[12] = {q.fSrcX, q.fSrcY, q.fSrcZ, q.fDetSX, q.fDetSY, q.fDetSZ,
                                  q.fDetUX, q.fDetUY, q.fDetUZ, q.fDetVX, q.fDetVY, q.fDetVZ};
            std::memcpy(&v[12 * (size_t)i], a, sizeof a);
        }
        return true;
    }
    if (const astra::SPar3DProjection *p = g.getParallel()) {
        kind = ASTRA_AMD_GEOM_PAR3D;
        v.resize(12 * (size_t)n);
        for (unsigned i = 0; i < n; ++i) {
            const astra::SPar3DProjection &q = p[i];
            const double a[12] = {q.fRayX, q.fRayY, q.fRayZ, q.fDetSX, q.fDetSY, q.fDetSZ,
                                  q.fDetUX, q.fDetUY, q.fDetUZ, q.fDetVX, q.fDetVY, q.fDetVZ};
            std::memcpy(&v[12 * (size_t)i], a, sizeof a);
        }
        return true;
    }
    if (const astra::SCylConeProjection *p = g.getCylCone()) {
        kind = ASTRA_AMD_GEOM_CYLCONE;  // translated only (GeometryUtil3D.cpp:597-616)
        v.resize(13 * (size_t)n);
        for (unsigned i = 0; i < n; ++i) {
            const astra::SCylConeProjection &q = p[i];
            const double a[13] = {q.fSrcX,  q.fSrcY,  q.fSrcZ,  q.fDetCX, q.fDetCY,
                                  q.fDetCZ, q.fDetUX, q.fDetUY, q.fDetUZ, q.fDetVX,
                                  q.fDetVY, q.fDetVZ, q.fDetR};
            std::memcpy(&v[13 * (size_t)i], a, sizeof a);
        }
        return true;
    }
    return false;
}

// One geometry handle per host thread, rebuilt only when the geometry
// changes: an iterative algorithm calls FP and BP on the same geometry
// every iteration, and a handle holds the per-angle tables on the device.
struct CachedGeom {
    int kind = -1;
    astra::SDimensions3D dims;
    float scale[3] = {0.f, 0.f, 0.f};
    int device = -1;
    std::vector<double> vecs;
    astra_amd_geom3d h = nullptr;
    ~CachedGeom() {
        if (h) astra_amd_geom3d_destroy(h);
    }
};

astra_amd_geom3d geometry(const astra::Geometry3DParameters &g, const astra::SVolScale3D &vs) {
    thread_local CachedGeom c;
    int kind = -1;
    std::vector<double> v;
    if (!vectors(g, kind, v)) {
        ASTRA_ERROR("astra_amd seam: unsupported projection geometry");
        return nullptr;
    }
    const astra::SDimensions3D &d = g.getDims();
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    const float scale[3] = {vs.fX, vs.fY, vs.fZ};
    const bool same = c.h && c.kind == kind && c.device == dev && c.dims.iVolX == d.iVolX &&
                      c.dims.iVolY == d.iVolY && c.dims.iVolZ == d.iVolZ &&
                      c.dims.iProjAngles == d.iProjAngles && c.dims.iProjU == d.iProjU &&
                      c.dims.iProjV == d.iProjV && std::memcmp(c.scale, scale, sizeof scale) == 0 &&
                      c.vecs == v;
    if (same) return c.h;
    if (c.h) astra_amd_geom3d_destroy(c.h);
    c.h = nullptr;
    astra_amd_geom3d h = nullptr;
    if (astra_amd_geom3d_create(kind, d.iVolX, d.iVolY, d.iVolZ, d.iProjAngles, d.iProjU,
                                d.iProjV, v.data(), scale, &h) != ASTRA_AMD_OK) {
        ASTRA_ERROR("astra_amd: %s", astra_amd_last_error());
        return nullptr;
    }
    c.kind = kind;
    c.dims = d;
    std::memcpy(c.scale, scale, sizeof scale);
    c.device = dev;
    c.vecs.swap(v);
    c.h = h;
    return h;
}

int kernel_of(astra::Cuda3DProjectionKernel k) {
    switch (k) {
    case astra::ker3d_default: return ASTRA_AMD_KER3D_DEFAULT;
    case astra::ker3d_sum_square_weights: return ASTRA_AMD_KER3D_SUM_SQUARE_WEIGHTS;
    case astra::ker3d_fdk_weighting: return ASTRA_AMD_KER3D_FDK_WEIGHTING;
    case astra::ker3d_2d_weighting: return ASTRA_AMD_KER3D_2D_WEIGHTING;
    case astra::ker3d_matched_bp: return ASTRA_AMD_KER3D_MATCHED_BP;
    }
    return -1;
}

// A C-contiguous view of pitched device data: the data itself when the
// pitch equals the row width, else a copy made on construction and copied
// back by write_back().
class Contiguous {
  public:
    Contiguous(const cudaPitchedPtr &p, size_t width, size_t rows)
        : src_(p), width_(width * sizeof(float)), rows_(rows) {
        if (src_.pitch == width_) {
            data_ = static_cast<float *>(src_.ptr);
            return;
        }
        if (hipMalloc(reinterpret_cast<void **>(&copy_), width_ * rows_) != hipSuccess) return;
        if (hipMemcpy2D(copy_, width_, src_.ptr, src_.pitch, width_, rows_,
                        hipMemcpyDeviceToDevice) != hipSuccess)
            return;
        data_ = copy_;
    }
    ~Contiguous() {
        if (copy_) (void)hipFree(copy_);
    }
    Contiguous(const Contiguous &) = delete;
    Contiguous &operator=(const Contiguous &) = delete;
    bool ok() const { return data_ != nullptr; }
    float *data() const { return data_; }
    bool write_back() const {
        if (!copy_) return data_ != nullptr;
        return hipMemcpy2D(src_.ptr, src_.pitch, copy_, width_, width_, rows_,
                           hipMemcpyDeviceToDevice) == hipSuccess;
    }

  private:
    cudaPitchedPtr src_;
    size_t width_, rows_;
    float *data_ = nullptr;
    float *copy_ = nullptr;
};

// The reference's functions return after the work is done (cuda/3d/*.cu
// synchronise their streams), so the seam waits for the null stream.
bool finish(int rc) {
    if (rc == ASTRA_AMD_OK) rc = astra_amd_stream_sync(nullptr);
    if (rc != ASTRA_AMD_OK) ASTRA_ERROR("astra_amd: %s", astra_amd_last_error());
    return rc == ASTRA_AMD_OK;
}

}  // namespace

namespace astraCUDA3d {

bool FP(astra::CData3D *projData, const astra::CData3D *volData,
        const astra::Geometry3DParameters &geom, SProjectorParams3D params) {
    cudaPitchedPtr vp, pp;
    if (!pitched(volData, vp) || !pitched(projData, pp)) return false;
    astra_amd_geom3d h = geometry(geom, params.volScale);
    const int kernel = kernel_of(params.projKernel);
    if (!h || kernel < 0) return false;
    // the FP takes the default kernel for fdk / 2d weighting / matched_bp
    // (cuda/3d/mem3d.cu:360-388)
    astra_amd_params3d q = {params.iRaysPerDetDim, params.iRaysPerVoxelDim, params.fOutputScale,
                            kernel == ASTRA_AMD_KER3D_SUM_SQUARE_WEIGHTS
                                ? ASTRA_AMD_KER3D_SUM_SQUARE_WEIGHTS
                                : ASTRA_AMD_KER3D_DEFAULT};
    return finish(astra_amd_fp3d_pitched(h, static_cast<const float *>(vp.ptr),
                                         vp.pitch / sizeof(float), static_cast<float *>(pp.ptr),
                                         pp.pitch / sizeof(float), &q, ASTRA_AMD_MODE_ADD,
                                         nullptr));
}

bool BP(const astra::CData3D *projData, astra::CData3D *volData,
        const astra::Geometry3DParameters &geom, SProjectorParams3D params) {
    cudaPitchedPtr pp, vp;
    if (!pitched(projData, pp) || !pitched(volData, vp)) return false;
    astra_amd_geom3d h = geometry(geom, params.volScale);
    const int kernel = kernel_of(params.projKernel);
    if (!h || kernel < 0) return false;
    astra_amd_params3d q = {params.iRaysPerDetDim, params.iRaysPerVoxelDim, params.fOutputScale,
                            kernel};
    return finish(astra_amd_bp3d_pitched(h, static_cast<const float *>(pp.ptr),
                                         pp.pitch / sizeof(float), static_cast<float *>(vp.ptr),
                                         vp.pitch / sizeof(float), &q, ASTRA_AMD_MODE_ADD,
                                         nullptr));
}

bool FDK(astra::CFloat32ProjectionData3D *projData, astra::CFloat32VolumeData3D *volData,
         bool bShortScan, const astra::SFilterConfig &cfg, float fOutputScale) {
    cudaPitchedPtr pp, vp;
    if (!pitched(projData, pp) || !pitched(volData, vp)) return false;
    const astra::Geometry3DParameters geom =
        astra::convertAstraGeometry(&volData->getGeometry(), &projData->getGeometry());
    if (!geom.isCone()) return false;  // as cuda/3d/mem3d.cu:506-507
    const astra::SDimensions3D &d = geom.getDims();
    astra_amd_geom3d h = geometry(geom, geom.getVolScale());
    if (!h) return false;

    // filter handle (src/Filters.cpp genFilter semantics, built by the library)
    astra_amd_filter f = nullptr;
    int rc = ASTRA_AMD_OK;
    const unsigned n_det = d.iProjU;
    unsigned n_fft = 1;
    while (n_fft < 2 * n_det) n_fft *= 2;  // calcNextPowerOfTwo (src/Filters.cpp:662-669)
    const unsigned n_bins = n_fft / 2 + 1;
    switch (cfg.m_eType) {
    case astra::FILTER_NONE:
        rc = astra_amd_filter_create(n_det, nullptr, 0, &f);
        break;
    case astra::FILTER_PROJECTION:
    case astra::FILTER_SINOGRAM:
    case astra::FILTER_RPROJECTION:
    case astra::FILTER_RSINOGRAM: {
        const bool per_angle =
            cfg.m_eType == astra::FILTER_SINOGRAM || cfg.m_eType == astra::FILTER_RSINOGRAM;
        const int real = cfg.m_eType == astra::FILTER_RPROJECTION ||
                         cfg.m_eType == astra::FILTER_RSINOGRAM;
        rc = astra_amd_filter_create_custom(n_det, per_angle ? cfg.m_iCustomFilterHeight : 1, real,
                                            cfg.m_pfCustomFilter.data(), cfg.m_iCustomFilterWidth,
                                            &f);
        break;
    }
    case astra::FILTER_ERROR:
        return false;
    default: {  // the windows FILTER_RAMLAK .. FILTER_PARZEN, in E_FBPFILTER order
        std::vector<float> spec(n_bins);
        rc = astra_amd_filter_spectrum(int(cfg.m_eType) - int(astra::FILTER_RAMLAK), n_det, cfg.m_fD,
                                       cfg.m_fParameter, spec.data(), n_bins);
        if (rc == ASTRA_AMD_OK) rc = astra_amd_filter_create(n_det, spec.data(), n_bins, &f);
    }
    }
    if (rc != ASTRA_AMD_OK) {
        ASTRA_ERROR("astra_amd: %s", astra_amd_last_error());
        return false;
    }
    // astra_amd_fdk takes C-contiguous data; cudaMalloc3D storage is
    // usually padded, so such data goes through contiguous copies (the
    // projections come back filtered, as the reference leaves them)
    Contiguous P(pp, d.iProjU, (size_t)d.iProjAngles * d.iProjV);
    Contiguous V(vp, d.iVolX, (size_t)d.iVolY * d.iVolZ);
    if (P.ok() && V.ok()) {
        // pre-weight + filter + FDK-weighted BP, accumulated like the
        // reference's ConeBP (the caller zeroes the volume for MODE_SET)
        rc = astra_amd_fdk(h, P.data(), V.data(), fOutputScale, bShortScan ? 1 : 0, f,
                           ASTRA_AMD_MODE_ADD, nullptr);
    } else {
        rc = ASTRA_AMD_ERR_ALLOC;
    }
    bool ok = finish(rc) && P.write_back() && V.write_back();
    if (!ok && rc == ASTRA_AMD_ERR_ALLOC) ASTRA_ERROR("astra_amd seam: FDK copy failed");
    astra_amd_filter_destroy(f);
    return ok;
}

}  // namespace astraCUDA3d
