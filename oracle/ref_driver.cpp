// C entry points over the REFERENCE CPU library (CParallelBeamLinearKernel-
// Projector2D, CParallelBeamStripKernelProjector2D, CSirtAlgorithm,
// CForward/CBackProjectionAlgorithm, the
// Shepp-Logan generators), built from the reference sources by
// oracle/Makefile into oracle/_ref/.
//
// TEST INFRASTRUCTURE ONLY: used by tests/ (golden checks of the C
// restatement) and by bench.py's cpu_baseline leg.  Never part of the product.
#include <cstring>
#include <vector>

#include "astra/BackProjectionAlgorithm.h"
#include "astra/Data2D.h"
#include "astra/ForwardProjectionAlgorithm.h"
#include "astra/ParallelBeamLinearKernelProjector2D.h"
#include "astra/ParallelBeamStripKernelProjector2D.h"
#include "astra/Data3D.h"
#include "astra/ParallelProjectionGeometry2D.h"
#include "astra/SheppLogan.h"
#include "astra/SirtAlgorithm.h"
#include "astra/VolumeGeometry2D.h"
#include "astra/VolumeGeometry3D.h"

using namespace astra;

namespace {
template <class Projector>
struct SetupT {
    CParallelProjectionGeometry2D pg;
    CVolumeGeometry2D vg;
    Projector proj;
    CFloat32ProjectionData2D sino;
    CFloat32VolumeData2D vol;
    SetupT(int rows, int cols, int na, int nd, float w, const float *angles)
        : pg(na, nd, w, std::vector<float32>(angles, angles + na)),
          vg(cols, rows),
          proj(pg, vg),
          sino(pg, new CDataMemory<float32>((size_t)na * nd)),
          vol(vg, new CDataMemory<float32>((size_t)rows * cols)) {}
};
using Setup = SetupT<CParallelBeamLinearKernelProjector2D>;
using StripSetup = SetupT<CParallelBeamStripKernelProjector2D>;
}  // namespace

extern "C" {

int ref_linear_fp(int rows, int cols, int na, int nd, float w, const float *angles,
                  const float *vol, float *sino) {
    Setup s(rows, cols, na, nd, w, angles);
    std::memcpy(s.vol.getFloat32Memory(), vol, sizeof(float) * (size_t)rows * cols);
    CForwardProjectionAlgorithm alg(&s.proj, &s.vol, &s.sino);
    if (!alg.isInitialized() || !alg.run(1)) return -1;
    std::memcpy(sino, s.sino.getFloat32Memory(), sizeof(float) * (size_t)na * nd);
    return 0;
}

int ref_linear_bp(int rows, int cols, int na, int nd, float w, const float *angles,
                  const float *sino, float *vol) {
    Setup s(rows, cols, na, nd, w, angles);
    std::memcpy(s.sino.getFloat32Memory(), sino, sizeof(float) * (size_t)na * nd);
    CBackProjectionAlgorithm alg(&s.proj, &s.sino, &s.vol);
    if (!alg.isInitialized() || !alg.run(1)) return -1;
    std::memcpy(vol, s.vol.getFloat32Memory(), sizeof(float) * (size_t)rows * cols);
    return 0;
}

int ref_linear_sirt(int rows, int cols, int na, int nd, float w, const float *angles,
                    const float *sino, float *vol, int iters) {
    Setup s(rows, cols, na, nd, w, angles);
    std::memcpy(s.sino.getFloat32Memory(), sino, sizeof(float) * (size_t)na * nd);
    std::memcpy(s.vol.getFloat32Memory(), vol, sizeof(float) * (size_t)rows * cols);
    CSirtAlgorithm alg;
    if (!alg.initialize(&s.proj, &s.sino, &s.vol) || !alg.run(iters)) return -1;
    std::memcpy(vol, s.vol.getFloat32Memory(), sizeof(float) * (size_t)rows * cols);
    return 0;
}

int ref_strip_fp(int rows, int cols, int na, int nd, float w, const float *angles,
                 const float *vol, float *sino) {
    StripSetup s(rows, cols, na, nd, w, angles);
    std::memcpy(s.vol.getFloat32Memory(), vol, sizeof(float) * (size_t)rows * cols);
    CForwardProjectionAlgorithm alg(&s.proj, &s.vol, &s.sino);
    if (!alg.isInitialized() || !alg.run(1)) return -1;
    std::memcpy(sino, s.sino.getFloat32Memory(), sizeof(float) * (size_t)na * nd);
    return 0;
}

int ref_strip_bp(int rows, int cols, int na, int nd, float w, const float *angles,
                 const float *sino, float *vol) {
    StripSetup s(rows, cols, na, nd, w, angles);
    std::memcpy(s.sino.getFloat32Memory(), sino, sizeof(float) * (size_t)na * nd);
    CBackProjectionAlgorithm alg(&s.proj, &s.sino, &s.vol);
    if (!alg.isInitialized() || !alg.run(1)) return -1;
    std::memcpy(vol, s.vol.getFloat32Memory(), sizeof(float) * (size_t)rows * cols);
    return 0;
}

int ref_shepp_logan_2d(int rows, int cols, int modified, float *out) {
    CVolumeGeometry2D vg(cols, rows);
    CFloat32VolumeData2D vol(vg, new CDataMemory<float32>((size_t)rows * cols));
    generateSheppLogan(&vol, modified != 0);
    std::memcpy(out, vol.getFloat32Memory(), sizeof(float) * (size_t)rows * cols);
    return 0;
}

int ref_shepp_logan_3d(int slices, int rows, int cols, int modified, float *out) {
    CVolumeGeometry3D vg(cols, rows, slices);
    CFloat32VolumeData3D *vol = createCFloat32VolumeData3DMemory(vg);
    if (!vol) return -1;
    generateSheppLogan3D(vol, modified != 0);
    std::memcpy(out, vol->getFloat32Memory(), sizeof(float) * (size_t)slices * rows * cols);
    delete vol;
    return 0;
}

}  // extern "C"
