// 2D projectors for gfx950: parallel / fan Joseph forward projection,
// voxel-driven back projection (plain and FBP-weighted), and the exact
// transpose of the forward projection.
//
// Data layout in HBM
//   user volume     [rows][cols] (row axis is world -y)
//   user sinogram   [angle][det]
//   FP input copies rows-as-slices  [h][w+4]   (vertical rays)
//                   cols-as-slices  [w][h+4]   (horizontal rays, transposed)
//   BP input copy   [angle][det+4]
// Copies carry a 2-texel zero border so that clamped coordinates read
// exactly the border-0 texture values without bounds checks.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <mutex>
#include <vector>

#include "proj2d_dev.hpp"

namespace amx {

// ------------------------------------------------------------ prep -------
// vol [h][w] -> rows copy [h][w+4] and cols copy [w][h+4] in one pass.
// grid = (ceil((w+4)/64), ceil((h+4)/64)).
__global__ __launch_bounds__(256) void vol2d_copies(const float *__restrict__ vol,
                                                    float *__restrict__ rows,
                                                    float *__restrict__ cols, unsigned w,
                                                    unsigned h) {
    __shared__ float tile[64][65];
    vol2d_copy_tile(vol, rows, cols, w, h, blockIdx.x * 64, blockIdx.y * 64, tile);
}

// sino [a][d] -> [a][d+4]; one workgroup per angle row
__global__ __launch_bounds__(256) void sino_pad(const float *__restrict__ sino,
                                                float *__restrict__ sp, unsigned nd) {
    const unsigned pd = nd + 2 * kPad;
    const float *src = sino + (size_t)blockIdx.x * nd;
    float *dst = sp + (size_t)blockIdx.x * pd;
    for (unsigned i = threadIdx.x; i < pd; i += blockDim.x) {
        int d = (int)i - kPad;
        dst[i] = (d >= 0 && d < (int)nd) ? src[d] : 0.f;
    }
}

// ------------------------------------------------------------- FP --------
struct FP2DArgs {
    const float *rows;  // [h][w+4]
    const float *cols;  // [w][h+4]
    unsigned w, h, na, nd;
    const FP2DAngle *ang;
    float osc;
    unsigned rays;  // rays per detector (DetectorSuperSampling)
    float *out;
    int mode;
    const float *rb, *rw;
    int padded_out;  // write [a][d+4]
};

// One thread per (detector, angle); 64 detectors x 4 angles per group.
template <bool FAN>
__global__ __launch_bounds__(256) void fp2d_kernel(FP2DArgs A) {
    const int det = blockIdx.x * 64 + (threadIdx.x & 63);
    const int angle = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (det >= (int)A.nd || angle >= (int)A.na) return;
    const float val =
        fp2d_value<FAN>(A.ang[angle], det, A.rays, A.w, A.h, A.rows, A.cols) * A.osc;
    const size_t idx = (size_t)angle * A.nd + det;
    const size_t oidx = A.padded_out ? (size_t)angle * (A.nd + 2 * kPad) + det + kPad : idx;
    store_fp(A.out, oidx, val, A.mode, A.rb, idx, A.rw);
}

// The same for rays = 1 with two neighbouring detectors per thread
// (fp2d_value_pair: one 16-byte load per slice serves both rays): 128
// detectors x 4 angles per group.  Equal to fp2d_kernel's values.
template <bool FAN>
__global__ __launch_bounds__(256) void fp2d_pair_kernel(FP2DArgs A) {
    const int det = blockIdx.x * 128 + 2 * (threadIdx.x & 63);
    const int angle = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (det >= (int)A.nd || angle >= (int)A.na) return;
    const float2 v = fp2d_value_pair<FAN>(A.ang[angle], det, A.w, A.h, A.rows, A.cols);
    for (int k = 0; k < 2; ++k) {
        const int d = det + k;
        if (d >= (int)A.nd) break;
        const size_t idx = (size_t)angle * A.nd + d;
        const size_t oidx = A.padded_out ? (size_t)angle * (A.nd + 2 * kPad) + d + kPad : idx;
        store_fp(A.out, oidx, (k ? v.y : v.x) * A.osc, A.mode, A.rb, idx, A.rw);
    }
}

// ------------------------------------------------------------- BP --------
struct BP2DArgs {
    const float *sp;  // [a][nd+4]
    unsigned w, h, na, nd;
    const BP2DAngle *ang;
    float osc;
    float *vol;
    int mode;
    const float *uw;
    float lo, hi;
};

// One thread per column x and 4 consecutive rows; 64 x 16 pixels per group.
// PR: rows in pairs sharing their loads (bp2d_column).
template <int KIND, bool PR = false>
__global__ __launch_bounds__(256) void bp2d_kernel(BP2DArgs A) {
    constexpr int PY = 4;
    const int X = blockIdx.x * 64 + (threadIdx.x & 63);
    const int Y0 = (blockIdx.y * 4 + (threadIdx.x >> 6)) * PY;
    if (X >= (int)A.w || Y0 >= (int)A.h) return;
    float acc[PY];
    bp2d_column<KIND, PY, PR>(X, Y0, A.w, A.h, A.na, A.nd, A.sp, A.ang, acc);
    store_bp_rows<PY>(A.vol, (size_t)Y0 * A.w + X, A.w, min(PY, (int)A.h - Y0), acc, A.osc, A.mode,
                      A.uw, A.lo, A.hi);
}

// The same with the sinogram footprints staged in LDS (bp2d_tile_lds):
// equal floats, taps from LDS instead of the texture-address path.
template <int KIND>
__global__ __launch_bounds__(256) void bp2d_lds_kernel(BP2DArgs A) {
    constexpr int PY = 4;
    __shared__ Bp2Lds L;
    const int X0 = blockIdx.x * 64, Yt0 = blockIdx.y * 4 * PY;
    float acc[PY];
    bp2d_tile_lds<KIND, PY>(X0, Yt0, A.w, A.h, A.na, A.nd, A.sp, A.ang, acc, L);
    const int X = X0 + (threadIdx.x & 63);
    const int Y0 = Yt0 + (threadIdx.x >> 6) * PY;
    if (X >= (int)A.w || Y0 >= (int)A.h) return;
    store_bp_rows<PY>(A.vol, (size_t)Y0 * A.w + X, A.w, min(PY, (int)A.h - Y0), acc, A.osc, A.mode,
                      A.uw, A.lo, A.hi);
}

// Supersampled BP (devBP_SS / devFanBP_SS, cuda/2d/par_bp.cu:90-137,
// cuda/2d/fan_bp.cu:106-161): R x R sub-pixels per pixel at offsets
// (k + 0.5) / R - 0.5, averaged.  KIND 0: parallel (angle weight), 1: fan
// (weight fr).  One thread per pixel; 64 x 4 pixels per group.
template <int KIND>
__global__ __launch_bounds__(256) void bp2d_ss_kernel(BP2DArgs A, int R) {
    const int X = blockIdx.x * 64 + (threadIdx.x & 63);
    const int Y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (X >= (int)A.w || Y >= (int)A.h) return;
    const float sub = 1.0f / (float)R;
    const float fx0 = X - 0.5f * A.w + 0.5f * sub;
    const float fy0 = -(Y - 0.5f * A.h + 0.5f * sub);  // world y of the first sub-row
    const unsigned pitch = A.nd + 2 * kPad;
    const float hi = (float)A.nd + 2.5f;
    float acc = 0.f;
    const float *row = A.sp;
    for (unsigned a = 0; a < A.na; ++a, row += pitch) {
        const BP2DAngle c = A.ang[a];
        for (int i = 0; i < R; ++i) {
            const float fx = fx0 + i * sub;
            for (int j = 0; j < R; ++j) {
                const float fy = fy0 - j * sub;
                float num = fmaf(fy, c.num[2], fmaf(fx, c.num[1], c.num[0]));
                float fr = 1.f;
                if (KIND != 0) fr = __builtin_amdgcn_rcpf(fmaf(fy, c.den[2], fmaf(fx, c.den[1], c.den[0])));
                float t = clamp_pad(num * fr, hi);
                float f = floorf(t);
                float2 q = ld2b(row, (unsigned)f << 2);
                acc = fmaf((KIND == 0) ? c.w : fr, lerp(q.x, q.y, t - f), acc);
            }
        }
    }
    store_bp(A.vol, (size_t)Y * A.w + X, acc * A.osc / (float)(R * R), A.mode, A.uw, A.lo, A.hi);
}

// --------------------------------------------------- matched BP ----------
// Exact transpose of fp2d_kernel as a gather (deterministic, no atomics).
// The FP takes, per (sub-)ray of lattice index L = R det + k and slice s of
// the ray's orientation class, the linear sample at padded coordinate
// t = a s + c (fp32) of the rows copy (vertical rays: slice = row, t along
// the columns) or of the columns copy (horizontal rays: slice = column, t
// along the rows).  Pixel (X, Y) therefore receives, per angle,
//     sum over vertical rays   of y' (1 - |a Y + c - (X + 2)|)+
//   + sum over horizontal rays of y' (1 - |a X + c - (Y + 2)|)+,
// y' = y * out_scale * distance weight / R.  Candidates are the lattice
// points between the images of the two neighbouring pixel centres along the
// class's in-slice axis under the BP's detector map (padded texel
// coordinate t_pad = num / den; lattice = R (t_pad - 1.5) - 0.5), widened by
// 0.05 for fp32 rounding; extra candidates get weight 0.
//
// matched2d_rays_kernel: per (angle, lattice index) the ray's fp32 (a, c),
// y' and its class (1 = vertical).
template <bool FAN>
__global__ __launch_bounds__(256) void matched2d_rays_kernel(FP2DArgs A, const float *__restrict__ sino,
                                                             float4 *__restrict__ rays) {
    const unsigned NL = A.nd * A.rays;
    const size_t n = (size_t)NL * A.na;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const unsigned angle = (unsigned)(i / NL), L = (unsigned)(i - (size_t)angle * NL);
        const int det = (int)(L / A.rays), k = (int)(L % A.rays);
        const FP2DAngle &G = A.ang[angle];
        const bool vertical = ray_vertical<FAN>(G, det);
        const double fdet = det + (0.5 + k) / (double)A.rays;
        const Ray2D R = ray2d<FAN>(G, fdet, vertical, A.w, A.h);
        const float y0 = sino[(size_t)angle * A.nd + det] * (A.osc / (float)A.rays);
        rays[i] = make_float4(R.a, R.c, y0 * R.scale, vertical ? 1.f : 0.f);
    }
}

// lattice coordinate of world point (x, y) under the BP's detector map
template <bool FAN>
__device__ __forceinline__ float lattice2d(const BP2DAngle &c, float x, float y, float R) {
    const float num = fmaf(y, c.num[2], fmaf(x, c.num[1], c.num[0]));
    const float t = FAN ? num / fmaf(y, c.den[2], fmaf(x, c.den[1], c.den[0])) : num;
    return fmaf(R, t - 1.5f, -0.5f);
}

// candidates [l0, l1] of one orientation class for slice `sf` and in-slice
// padded target J, weighted into acc
__device__ __forceinline__ float gather2d_class(const float4 *__restrict__ ra, float la, float lb,
                                                float NLm, float want, float sf, float J) {
    constexpr float kEps = 0.05f;
    const int l0 = (int)fminf(fmaxf(ceilf(fminf(la, lb) - kEps), 0.f), NLm + 1.f);
    const int l1 = (int)fmaxf(fminf(floorf(fmaxf(la, lb) + kEps), NLm), -1.f);
    float acc = 0.f;
    for (int L = l0; L <= l1; ++L) {
        const float4 r = ra[L];
        const float e = tent(fmaf(r.x, sf, r.y) - J);
        acc = fmaf(r.w == want ? r.z : 0.f, e, acc);
    }
    return acc;
}

struct Matched2DArgs {
    const float4 *rays;      // [angle][nd * R]
    const BP2DAngle *bpa;    // detector map per angle (num / den)
    const FP2DAngle *fpa;    // per-angle class (parallel beams)
    unsigned w, h, na, nd, R;
    float *vol;
    int mode;
};

// One thread per pixel, 64 x 4 pixels per group, all angles.
template <bool FAN>
__global__ __launch_bounds__(256) void matched2d_gather(Matched2DArgs M) {
    const int X = blockIdx.x * 64 + (threadIdx.x & 63);
    const int Y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (X >= (int)M.w || Y >= (int)M.h) return;
    const float x = X - 0.5f * M.w + 0.5f, y = -(Y - 0.5f * M.h + 0.5f);
    const float Rf = (float)M.R;
    const unsigned NL = M.nd * M.R;
    const float NLm = (float)NL - 1.f;
    const float Jc = (float)(X + kPad), Jr = (float)(Y + kPad);
    float acc = 0.f;
    for (unsigned a = 0; a < M.na; ++a) {
        const BP2DAngle c = M.bpa[a];
        const float4 *ra = M.rays + (size_t)a * NL;
        // vertical rays: slice Y, neighbours X -/+ 1 along the columns
        if (FAN || M.fpa[a].vertical)
            acc += gather2d_class(ra, lattice2d<FAN>(c, x - 1.f, y, Rf),
                                  lattice2d<FAN>(c, x + 1.f, y, Rf), NLm, 1.f, (float)Y, Jc);
        // horizontal rays: slice X, neighbours Y -/+ 1 along the rows
        if (FAN || !M.fpa[a].vertical)
            acc += gather2d_class(ra, lattice2d<FAN>(c, x, y + 1.f, Rf),
                                  lattice2d<FAN>(c, x, y - 1.f, Rf), NLm, 0.f, (float)X, Jr);
    }
    float *o = M.vol + (size_t)Y * M.w + X;
    if (M.mode == OUT_SET)
        *o = acc;
    else
        *o += acc;
}

// ------------------------------------------------------------ launch -----
namespace {
size_t rows_elems(const Geom2D &g) { return (size_t)g.h * (g.w + 2 * kPad); }
size_t cols_elems(const Geom2D &g) { return (size_t)g.w * (g.h + 2 * kPad); }
}  // namespace

size_t padded_sino_elems(const Geom2D &g) { return (size_t)g.na * (g.nd + 2 * kPad); }

void sino_to_padded(const Geom2D &g, const float *sino, float *sp, hipStream_t s) {
    sino_pad<<<g.na, 256, 0, s>>>(sino, sp, g.nd);
    AMX_LAUNCHED();
}

void fp2d_launch(const Geom2D &g, const float *vol, float *sino, float out_scale, int mode,
                 const float *resid_b, const float *resid_w, bool padded_out, unsigned rays,
                 hipStream_t s) {
    fp2d_launch_range(g, vol, sino, out_scale, mode, resid_b, resid_w, padded_out, rays, 0, g.na,
                      s);
}

void fp2d_launch_range(const Geom2D &g, const float *vol, float *sino, float out_scale, int mode,
                       const float *resid_b, const float *resid_w, bool padded_out, unsigned rays,
                       unsigned a0, unsigned na_sub, hipStream_t s) {
    AMX_REQUIRE(rays >= 1 && rays <= 64, "fp2d: rays per detector must be in [1, 64]");
    AMX_REQUIRE(a0 + na_sub <= g.na && na_sub > 0, "fp2d: angle range out of bounds");
    Scratch copies(fp2d_copies_elems(g) * sizeof(float), s);
    fp2d_run(g, vol, sino, out_scale, mode, resid_b, resid_w, padded_out, rays, a0, na_sub,
             copies.as<float>(), true, s);
}

size_t fp2d_copies_elems(const Geom2D &g) { return rows_elems(g) + cols_elems(g); }

// FP through the caller's copy buffer `copies` (fp2d_copies_elems(g) floats,
// every element written here before it is read).  `prof` = false records no
// timing events.
void fp2d_run(const Geom2D &g, const float *vol, float *sino, float out_scale, int mode,
              const float *resid_b, const float *resid_w, bool padded_out, unsigned rays,
              unsigned a0, unsigned na_sub, float *copies, bool prof, hipStream_t s) {
    AMX_REQUIRE(rays >= 1 && rays <= 64, "fp2d: rays per detector must be in [1, 64]");
    AMX_REQUIRE(a0 + na_sub <= g.na && na_sub > 0, "fp2d: angle range out of bounds");
    float *rows = copies;
    float *cols = rows + rows_elems(g);
    dim3 pg((g.w + 2 * kPad + 63) / 64, (g.h + 2 * kPad + 63) / 64);
    vol2d_copies<<<pg, 256, 0, s>>>(vol, rows, cols, g.w, g.h);
    AMX_LAUNCHED();
    FP2DArgs A{};
    A.rows = rows;
    A.cols = cols;
    A.w = g.w; A.h = g.h; A.na = g.na; A.nd = g.nd;
    // angles [a0, a0 + na_sub): offset the angle table and the row pointers
    const size_t out_pitch = padded_out ? g.nd + 2 * kPad : g.nd;
    A.ang = g.fp.as<FP2DAngle>() + a0;
    A.na = na_sub;
    A.osc = out_scale;
    A.rays = rays;
    A.out = sino + (size_t)a0 * out_pitch;
    A.mode = mode;
    A.rb = resid_b ? resid_b + (size_t)a0 * g.nd : nullptr;
    A.rw = resid_w ? resid_w + (size_t)a0 * out_pitch : nullptr;
    A.padded_out = padded_out ? 1 : 0;
    dim3 grid((g.nd + 63) / 64, (na_sub + 3) / 4);
    ProfToken pt{};
    if (prof)
        pt = prof_begin(g.kind == ASTRA_AMD_GEOM_FAN ? "fp2d_fan" : "fp2d_par",
                        (double)na_sub * g.nd * rays * std::max(g.w, g.h), s);
    if (rays == 1 && fp2d_pairs()) {
        dim3 pgrid((g.nd + 127) / 128, (na_sub + 3) / 4);
        if (g.kind == ASTRA_AMD_GEOM_FAN)
            fp2d_pair_kernel<true><<<pgrid, 256, 0, s>>>(A);
        else
            fp2d_pair_kernel<false><<<pgrid, 256, 0, s>>>(A);
    } else if (g.kind == ASTRA_AMD_GEOM_FAN)
        fp2d_kernel<true><<<grid, 256, 0, s>>>(A);
    else
        fp2d_kernel<false><<<grid, 256, 0, s>>>(A);
    AMX_LAUNCHED();
    if (prof) prof_end(pt, s);
}

// ---- FP that allocates nothing --------------------------------------------
// For callers that may not allocate device memory (code under stream
// capture, serving loops, harnesses that forbid hipMalloc): the volume
// copies live in this fixed array, one per GPU, held until the process
// ends.  Every call on a GPU shares that GPU's array, so calls that overlap
// on one GPU (two streams, two host threads) race; nothing in it needs to
// start at zero (fp2d_run writes every element it reads).
constexpr size_t kNoAllocFloats = size_t(16) << 20;  // 64 MiB
static __device__ __align__(256) float scratch[kNoAllocFloats];

namespace {
std::mutex g_scratch_mu;
std::vector<float *> g_scratch_ptr;  // per device: address of that GPU's array

// The array's device address on the current device.  The host-side symbol
// `scratch` is only a handle: it is never passed to a kernel or a copy.
float *noalloc_scratch() {
    int dev = -1;
    AMX_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((size_t)dev >= g_scratch_ptr.size()) g_scratch_ptr.resize(dev + 1, nullptr);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        AMX_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        AMX_REQUIRE(ptr != nullptr, "fp2d_noalloc: no device address for the scratch array");
        g_scratch_ptr[dev] = static_cast<float *>(ptr);
    }
    return g_scratch_ptr[dev];
}
}  // namespace

size_t fp2d_noalloc_max_floats() { return kNoAllocFloats; }

void fp2d_launch_noalloc(const Geom2D &g, const float *vol, float *sino, float out_scale,
                         int mode, unsigned rays, hipStream_t s) {
    const size_t need = fp2d_copies_elems(g);
    AMX_REQUIRE(need <= kNoAllocFloats,
                "fp2d_noalloc: a %u x %u volume needs %zu floats of scratch, more than the "
                "fixed %zu (64 MiB); use astra_amd_fp2d",
                g.w, g.h, need, kNoAllocFloats);
    fp2d_run(g, vol, sino, out_scale, mode, nullptr, nullptr, false, rays, 0, g.na,
             noalloc_scratch(), false, s);
}

// ASTRA_AMD_FP2D_PAIRS=1 projects two neighbouring detectors per thread,
// sharing their loads, instead of one (read at every launch; measured
// slower: 0.385 vs 0.208 s for the fan 1024^2 SIRT with the row pairs below)
bool fp2d_pairs() {
    const char *e = std::getenv("ASTRA_AMD_FP2D_PAIRS");
    return e && e[0] == '1';
}

// ASTRA_AMD_BP2D_PAIRS=1 loads the taps of two rows with one 16-byte load
// where they are close enough, instead of each row's separately (read at
// every launch; measured no faster)
bool bp2d_pairs() {
    const char *e = std::getenv("ASTRA_AMD_BP2D_PAIRS");
    return e && e[0] == '1';
}

// ASTRA_AMD_BP2D_LDS=1 takes the taps of the 2D BP from LDS-staged
// footprints instead of global memory (read at every launch)
bool bp2d_lds() {
    const char *e = std::getenv("ASTRA_AMD_BP2D_LDS");
    return e && e[0] == '1';
}

void bp2d_launch_padded(const Geom2D &g, const float *sp, float *vol, float out_scale, bool fbp,
                        int mode, const float *upd_w, float lo, float hi, hipStream_t s,
                        unsigned ss) {
    AMX_REQUIRE(ss >= 1 && ss <= 16, "bp2d: sub-pixels per dimension must be in [1, 16]");
    BP2DArgs A{};
    A.sp = sp;
    A.w = g.w; A.h = g.h; A.na = g.na; A.nd = g.nd;
    A.osc = out_scale;
    A.vol = vol;
    A.mode = mode;
    A.uw = upd_w;
    A.lo = lo;
    A.hi = hi;
    dim3 grid((g.w + 63) / 64, (g.h + 15) / 16);
    ProfToken pt = prof_begin(g.kind == ASTRA_AMD_GEOM_PAR ? "bp2d_par" : (fbp ? "bp2d_fbp" : "bp2d_fan"),
                              (double)g.w * g.h * g.na, s);
    if (ss > 1 && !fbp) {  // the FBP-weighted BP has no supersampled form
        A.ang = g.bp.as<BP2DAngle>();
        dim3 sgrid((g.w + 63) / 64, (g.h + 3) / 4);
        if (g.kind == ASTRA_AMD_GEOM_PAR)
            bp2d_ss_kernel<0><<<sgrid, 256, 0, s>>>(A, (int)ss);
        else
            bp2d_ss_kernel<1><<<sgrid, 256, 0, s>>>(A, (int)ss);
    } else {
        const int kind = g.kind == ASTRA_AMD_GEOM_PAR ? 0 : (fbp ? 2 : 1);
        A.ang = (kind == 2 ? g.bp_fbp : g.bp).as<BP2DAngle>();
        const bool pr = bp2d_pairs();
        if (bp2d_lds()) {
            if (kind == 0) bp2d_lds_kernel<0><<<grid, 256, 0, s>>>(A);
            else if (kind == 1) bp2d_lds_kernel<1><<<grid, 256, 0, s>>>(A);
            else bp2d_lds_kernel<2><<<grid, 256, 0, s>>>(A);
        } else if (pr) {
            if (kind == 0) bp2d_kernel<0, true><<<grid, 256, 0, s>>>(A);
            else if (kind == 1) bp2d_kernel<1, true><<<grid, 256, 0, s>>>(A);
            else bp2d_kernel<2, true><<<grid, 256, 0, s>>>(A);
        } else {
            if (kind == 0) bp2d_kernel<0><<<grid, 256, 0, s>>>(A);
            else if (kind == 1) bp2d_kernel<1><<<grid, 256, 0, s>>>(A);
            else bp2d_kernel<2><<<grid, 256, 0, s>>>(A);
        }
    }
    AMX_LAUNCHED();
    prof_end(pt, s);
}

void bp2d_sart_launch(const Geom2D &g, const float *sp_row, float *vol, unsigned angle,
                      float out_scale, const float *upd_w, float lo, float hi, hipStream_t s) {
    AMX_REQUIRE(angle < g.na, "bp2d_sart: angle out of range");
    BP2DArgs A{};
    A.sp = sp_row;
    A.w = g.w; A.h = g.h; A.na = 1; A.nd = g.nd;
    A.ang = g.bp.as<BP2DAngle>() + angle;
    A.osc = out_scale;
    A.vol = vol;
    A.mode = OUT_UPDATE;
    A.uw = upd_w;
    A.lo = lo;
    A.hi = hi;
    dim3 grid((g.w + 63) / 64, (g.h + 15) / 16);
    if (g.kind == ASTRA_AMD_GEOM_PAR)
        bp2d_kernel<3><<<grid, 256, 0, s>>>(A);
    else
        bp2d_kernel<4><<<grid, 256, 0, s>>>(A);
    AMX_LAUNCHED();
}

void bp2d_launch(const Geom2D &g, const float *sino, float *vol, float out_scale, bool fbp,
                 int mode, const float *upd_w, float lo, float hi, hipStream_t s, unsigned ss) {
    Scratch sp(padded_sino_elems(g) * sizeof(float), s);
    sino_to_padded(g, sino, sp.as<float>(), s);
    bp2d_launch_padded(g, sp.as<float>(), vol, out_scale, fbp, mode, upd_w, lo, hi, s, ss);
}

void bp2d_matched_launch(const Geom2D &g, const float *sino, float *vol, float out_scale,
                         int mode, hipStream_t s, unsigned rays) {
    AMX_REQUIRE(rays >= 1 && rays <= 64, "bp2d: rays per detector must be in [1, 64]");
    AMX_REQUIRE((size_t)g.nd * rays < (1u << 24), "bp2d: too many rays per angle");
    FP2DArgs A{};
    A.w = g.w; A.h = g.h; A.na = g.na; A.nd = g.nd;
    A.ang = g.fp.as<FP2DAngle>();
    A.osc = out_scale;
    A.rays = rays;
    const size_t n = (size_t)g.na * g.nd * rays;
    Scratch table(n * sizeof(float4), s);
    const unsigned rb = (unsigned)std::min<size_t>((n + 255) / 256, 1u << 20);
    Matched2DArgs M{};
    M.rays = table.as<float4>();
    M.bpa = g.bp.as<BP2DAngle>();
    M.fpa = g.fp.as<FP2DAngle>();
    M.w = g.w; M.h = g.h; M.na = g.na; M.nd = g.nd; M.R = rays;
    M.vol = vol;
    M.mode = mode;
    dim3 grid((g.w + 63) / 64, (g.h + 3) / 4);
    if (g.kind == ASTRA_AMD_GEOM_FAN) {
        matched2d_rays_kernel<true><<<rb, 256, 0, s>>>(A, sino, table.as<float4>());
        matched2d_gather<true><<<grid, 256, 0, s>>>(M);
    } else {
        matched2d_rays_kernel<false><<<rb, 256, 0, s>>>(A, sino, table.as<float4>());
        matched2d_gather<false><<<grid, 256, 0, s>>>(M);
    }
    AMX_LAUNCHED();
}

}  // namespace amx
