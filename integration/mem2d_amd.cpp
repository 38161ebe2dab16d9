// Reference-side binding of the 2D seam: new bodies for astraCUDA::FP and
// astraCUDA::BP (include/astra/cuda/2d/mem2d.h:77-83) that call
// libastra_amd through its C ABI (include/astra_amd.h).
//
// This file belongs to an ASTRA source tree, not to libastra_amd: a
// maintainer builds it in place of cuda/2d/mem2d.cu's FP / BP (the
// reference's HIP configuration, ASTRA_BUILDING_HIP) and links
// -lastra_amd.  The product never includes or builds it; a CPU test
// compiles it with -fsyntax-only against the reference headers.
//
// Replaced bodies: cuda/2d/mem2d.cu:224-261 (FP), :278-315 (BP).

#include "astra/cuda/gpu_runtime_wrapper.h"

#include "astra/cuda/2d/mem2d.h"
#include "astra/cuda/3d/mem3d_internal.h"
#include "astra/Data2D.h"
#include "astra/GeometryUtil2D.h"
#include "astra/Logging.h"

#include <cstring>
#include <vector>

#include "astra_amd.h"

namespace {

// Pitched device storage of a 2D data object (the same astraCUDA::CDataGPU
// storage class as in 3D, mem3d_internal.h:38-60); cudaArray-backed
// storage is refused, as cuda/2d/mem2d.cu:232-233 asserts.
bool pitched(const astra::CData2D *d, cudaPitchedPtr &out) {
    const astraCUDA::CDataGPU *s = dynamic_cast<const astraCUDA::CDataGPU *>(d->getStorage());
    if (!s || s->getArray() != nullptr) {
        ASTRA_ERROR("astra_amd seam: data object is not pitched GPU memory");
        return false;
    }
    out = s->getPtr();
    return out.ptr != nullptr;
}

// 6 floats per angle [ray or source, dS, u], already normalised by
// convertAstraGeometry (src/GeometryUtil2D.cpp:181-209).
bool vectors(const astra::Geometry2DParameters &g, int &kind, std::vector<float> &v) {
    const unsigned n = g.getDims().iProjAngles;
    v.resize(6 * (size_t)n);
    if (const astra::SParProjection *p = g.getParallel()) {
        kind = ASTRA_AMD_GEOM_PAR;
        for (unsigned i = 0; i < n; ++i) {
            const float a[6] = {p[i].fRayX, p[i].fRayY, p[i].fDetSX,
                                p[i].fDetSY, p[i].fDetUX, p[i].fDetUY};
            std::memcpy(&v[6 * (size_t)i], a, sizeof a);
        }
        return true;
    }
    if (const astra::SFanProjection *p = g.getFan()) {
        kind = ASTRA_AMD_GEOM_FAN;
        for (unsigned i = 0; i < n; ++i) {
            const float a[6] = {p[i].fSrcX, p[i].fSrcY, p[i].fDetSX,
                                p[i].fDetSY, p[i].fDetUX, p[i].fDetUY};
            std::memcpy(&v[6 * (size_t)i], a, sizeof a);
        }
        return true;
    }
    return false;
}

// One geometry handle per host thread, rebuilt only when the geometry
// changes (SIRT calls FP and BP on one geometry every iteration).
struct CachedGeom {
    int kind = -1;
    astra::SDimensions dims{};
    int device = -1;
    std::vector<float> vecs;
    astra_amd_geom2d h = nullptr;
    ~CachedGeom() {
        if (h) astra_amd_geom2d_destroy(h);
    }
};

astra_amd_geom2d geometry(const astra::Geometry2DParameters &g) {
    thread_local CachedGeom c;
    int kind = -1;
    std::vector<float> v;
    if (!vectors(g, kind, v)) {
        ASTRA_ERROR("astra_amd seam: unsupported projection geometry");
        return nullptr;
    }
    const astra::SDimensions &d = g.getDims();
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    if (c.h && c.kind == kind && c.device == dev && c.dims.iVolWidth == d.iVolWidth &&
        c.dims.iVolHeight == d.iVolHeight && c.dims.iProjAngles == d.iProjAngles &&
        c.dims.iProjDets == d.iProjDets && c.vecs == v)
        return c.h;
    if (c.h) astra_amd_geom2d_destroy(c.h);
    c.h = nullptr;
    astra_amd_geom2d h = nullptr;
    if (astra_amd_geom2d_create(kind, d.iVolWidth, d.iVolHeight, d.iProjAngles, d.iProjDets,
                                v.data(), &h) != ASTRA_AMD_OK) {
        ASTRA_ERROR("astra_amd: %s", astra_amd_last_error());
        return nullptr;
    }
    c.kind = kind;
    c.dims = d;
    c.device = dev;
    c.vecs.swap(v);
    c.h = h;
    return h;
}

// synchronous, as the reference's FP / BP
bool finish(int rc) {
    if (rc == ASTRA_AMD_OK) rc = astra_amd_stream_sync(nullptr);
    if (rc != ASTRA_AMD_OK) ASTRA_ERROR("astra_amd: %s", astra_amd_last_error());
    return rc == ASTRA_AMD_OK;
}

}  // namespace

namespace astraCUDA {

bool FP(astra::CData2D *projData, const astra::CData2D *volData,
        const astra::Geometry2DParameters &geom, SProjectorParams2D params) {
    cudaPitchedPtr vp, pp;
    if (!pitched(volData, vp) || !pitched(projData, pp)) return false;
    astra_amd_geom2d h = geometry(geom);
    if (!h) return false;
    const astra_amd_params2d q = {params.iRaysPerDet, params.iRaysPerPixelDim,
                                  params.fOutputScale};
    return finish(astra_amd_fp2d_pitched(h, static_cast<const float *>(vp.ptr),
                                         vp.pitch / sizeof(float), static_cast<float *>(pp.ptr),
                                         pp.pitch / sizeof(float), &q, ASTRA_AMD_MODE_ADD,
                                         nullptr));
}

bool BP(const astra::CData2D *projData, astra::CData2D *volData,
        const astra::Geometry2DParameters &geom, SProjectorParams2D params) {
    cudaPitchedPtr pp, vp;
    if (!pitched(projData, pp) || !pitched(volData, vp)) return false;
    astra_amd_geom2d h = geometry(geom);
    if (!h) return false;
    const astra_amd_params2d q = {params.iRaysPerDet, params.iRaysPerPixelDim,
                                  params.fOutputScale};
    return finish(astra_amd_bp2d_pitched(h, static_cast<const float *>(pp.ptr),
                                         pp.pitch / sizeof(float), static_cast<float *>(vp.ptr),
                                         vp.pitch / sizeof(float), &q, ASTRA_AMD_MODE_ADD,
                                         nullptr));
}

}  // namespace astraCUDA
