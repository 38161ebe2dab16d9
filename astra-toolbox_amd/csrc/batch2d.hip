// Batched 2D SIRT: thousands of 2D slices of uneven size, angle count and
// geometry (parallel / fan) reconstructed together.
//
// All slices live in packed device buffers.  Each pipeline stage (volume
// slice copies, residual forward projection, weighted back projection with
// the SIRT update) is ONE launch over the whole batch: a flattened table of
// work items (per-slice prefix sums) maps every workgroup to a (slice, tile)
// pair, so a 128^2 slice and a 1024^2 slice share the GPU in the same
// launch.  N iterations are captured once into a hipGraph (3N kernel nodes)
// and replayed as a single graph launch.  FP and BP tile shapes are chosen
// per batch by timing the candidates on the batch itself.
//
// The per-ray / per-pixel arithmetic is proj2d_dev.hpp, shared with the
// single-slice projectors, so a batch reproduces per-slice SIRT_CUDA runs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <limits>
#include <memory>
#include <mutex>

#include "proj2d_dev.hpp"

namespace amx {

struct SliceDesc {
    int kind;  // ASTRA_AMD_GEOM_PAR / _FAN
    unsigned w, h, na, nd;
    float osc;
    unsigned long long vol_off, sino_off, psino_off, rows_off, cols_off;
    unsigned ang_off;
    unsigned pad_;
};

struct BatchArgs {
    const SliceDesc *desc;
    unsigned n;
    const unsigned *prefix;  // [n+1] work items per slice (stage specific)
    const FP2DAngle *fpa;
    const BP2DAngle *bpa;
};

__device__ __forceinline__ unsigned find_slice(const unsigned *prefix, unsigned n, unsigned item) {
    unsigned lo = 0, hi = n;  // prefix[lo] <= item < prefix[hi]
    while (hi - lo > 1) {
        unsigned mid = (lo + hi) >> 1;
        if (prefix[mid] <= item)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// ---------------------------------------------------------------- prep ---
__global__ __launch_bounds__(256) void batch_prep(BatchArgs B, const float *__restrict__ vols,
                                                  float *__restrict__ rows,
                                                  float *__restrict__ cols) {
    __shared__ float tile[64][65];
    const unsigned item = blockIdx.x;
    const unsigned s = find_slice(B.prefix, B.n, item);
    const SliceDesc d = B.desc[s];
    const unsigned local = item - B.prefix[s];
    const unsigned tx = (d.w + 2 * kPad + 63) / 64;
    vol2d_copy_tile(vols + d.vol_off, rows + d.rows_off, cols + d.cols_off, d.w, d.h,
                    (local % tx) * 64, (local / tx) * 64, tile);
}

// sino rows -> padded [na][nd+4]; one workgroup per (slice, angle) row
__global__ __launch_bounds__(256) void batch_pad(BatchArgs B, const float *__restrict__ sinos,
                                                 float *__restrict__ psino) {
    const unsigned item = blockIdx.x;
    const unsigned s = find_slice(B.prefix, B.n, item);
    const SliceDesc d = B.desc[s];
    const unsigned a = item - B.prefix[s];
    const float *src = sinos + d.sino_off + (size_t)a * d.nd;
    float *dst = psino + d.psino_off + (size_t)a * (d.nd + 2 * kPad);
    for (unsigned i = threadIdx.x; i < d.nd + 2 * kPad; i += blockDim.x) {
        int j = (int)i - kPad;
        dst[i] = (j >= 0 && j < (int)d.nd) ? src[j] : 0.f;
    }
}

// padded [na][nd+4] -> packed sino rows (inverse of batch_pad)
__global__ __launch_bounds__(256) void batch_unpad(BatchArgs B, const float *__restrict__ psino,
                                                   float *__restrict__ sinos) {
    const unsigned item = blockIdx.x;
    const unsigned s = find_slice(B.prefix, B.n, item);
    const SliceDesc d = B.desc[s];
    const unsigned a = item - B.prefix[s];
    const float *src = psino + d.psino_off + (size_t)a * (d.nd + 2 * kPad) + kPad;
    float *dst = sinos + d.sino_off + (size_t)a * d.nd;
    for (unsigned i = threadIdx.x; i < d.nd; i += blockDim.x) dst[i] = src[i];
}

// ------------------------------------------------------------------ FP ---
// DB detectors x (256 / DB) angles per workgroup.  mode OUT_SET writes
// osc * Ax; OUT_RESID writes (b - osc * Ax) * w.  Output: padded sinogram.
template <int DB>
__global__ __launch_bounds__(256) void batch_fp(BatchArgs B, const float *__restrict__ rows,
                                                const float *__restrict__ cols,
                                                const float *__restrict__ sinos,
                                                const float *__restrict__ lw,
                                                float *__restrict__ out, int mode) {
    constexpr int AB = 256 / DB;
    const unsigned item = blockIdx.x;
    const unsigned s = find_slice(B.prefix, B.n, item);
    const SliceDesc d = B.desc[s];
    const unsigned local = item - B.prefix[s];
    const unsigned dblocks = (d.nd + DB - 1) / DB;
    const int det = (local % dblocks) * DB + (threadIdx.x % DB);
    const int angle = (local / dblocks) * AB + (threadIdx.x / DB);
    if (det >= (int)d.nd || angle >= (int)d.na) return;
    const FP2DAngle &G = B.fpa[d.ang_off + angle];
    const float *r = rows + d.rows_off, *c = cols + d.cols_off;
    float v = (d.kind == ASTRA_AMD_GEOM_FAN) ? fp2d_value<true>(G, det, 1, d.w, d.h, r, c)
                                             : fp2d_value<false>(G, det, 1, d.w, d.h, r, c);
    const size_t oidx = d.psino_off + (size_t)angle * (d.nd + 2 * kPad) + det + kPad;
    if (mode == OUT_SET) {
        out[oidx] = v * d.osc;
    } else {
        const size_t idx = d.sino_off + (size_t)angle * d.nd + det;
        out[oidx] = (sinos[idx] - v * d.osc) * lw[oidx];
    }
}

// The same with two neighbouring detectors per thread (fp2d_value_pair):
// DB detectors x (512 / DB) angles per workgroup.
template <int DB>
__global__ __launch_bounds__(256) void batch_fp_pair(BatchArgs B, const float *__restrict__ rows,
                                                     const float *__restrict__ cols,
                                                     const float *__restrict__ sinos,
                                                     const float *__restrict__ lw,
                                                     float *__restrict__ out, int mode) {
    constexpr int TD = DB / 2, AB = 256 / TD;
    const unsigned item = blockIdx.x;
    const unsigned s = find_slice(B.prefix, B.n, item);
    const SliceDesc d = B.desc[s];
    const unsigned local = item - B.prefix[s];
    const unsigned dblocks = (d.nd + DB - 1) / DB;
    const int det = (local % dblocks) * DB + 2 * (threadIdx.x % TD);
    const int angle = (local / dblocks) * AB + (threadIdx.x / TD);
    if (det >= (int)d.nd || angle >= (int)d.na) return;
    const FP2DAngle &G = B.fpa[d.ang_off + angle];
    const float *r = rows + d.rows_off, *c = cols + d.cols_off;
    const float2 v = (d.kind == ASTRA_AMD_GEOM_FAN) ? fp2d_value_pair<true>(G, det, d.w, d.h, r, c)
                                                     : fp2d_value_pair<false>(G, det, d.w, d.h, r, c);
    for (int k = 0; k < 2; ++k) {
        const int dd = det + k;
        if (dd >= (int)d.nd) break;
        const float val = k ? v.y : v.x;
        const size_t oidx = d.psino_off + (size_t)angle * (d.nd + 2 * kPad) + dd + kPad;
        if (mode == OUT_SET) {
            out[oidx] = val * d.osc;
        } else {
            const size_t idx = d.sino_off + (size_t)angle * d.nd + dd;
            out[oidx] = (sinos[idx] - val * d.osc) * lw[oidx];
        }
    }
}

// The same over 64 detectors x 16 angles per workgroup with the slice rows
// staged in LDS (fp2d_tile_lds).
__global__ __launch_bounds__(256) void batch_fp_lds(BatchArgs B, const float *__restrict__ rows,
                                                    const float *__restrict__ cols,
                                                    const float *__restrict__ sinos,
                                                    const float *__restrict__ lw,
                                                    float *__restrict__ out, int mode) {
    __shared__ Fp2Lds L;
    const unsigned item = blockIdx.x;
    const unsigned s = find_slice(B.prefix, B.n, item);
    const SliceDesc d = B.desc[s];
    const unsigned local = item - B.prefix[s];
    const unsigned dblocks = (d.nd + 63) / 64;
    const int det = (local % dblocks) * 64 + (threadIdx.x & 63);
    const int a0 = (local / dblocks) * 16 + (threadIdx.x >> 6) * kFp2Rays;
    const FP2DAngle *fpa = B.fpa + d.ang_off;
    const float *r = rows + d.rows_off, *c = cols + d.cols_off;
    float val[kFp2Rays];
    if (d.kind == ASTRA_AMD_GEOM_FAN)
        fp2d_tile_lds<true>(det, a0, d.na, d.nd, d.w, d.h, fpa, r, c, val, L);
    else
        fp2d_tile_lds<false>(det, a0, d.na, d.nd, d.w, d.h, fpa, r, c, val, L);
    if (det >= (int)d.nd) return;
#pragma unroll
    for (int j = 0; j < kFp2Rays; ++j) {
        const int angle = a0 + j;
        if (angle >= (int)d.na) break;
        const size_t oidx = d.psino_off + (size_t)angle * (d.nd + 2 * kPad) + det + kPad;
        if (mode == OUT_SET) {
            out[oidx] = val[j] * d.osc;
        } else {
            const size_t idx = d.sino_off + (size_t)angle * d.nd + det;
            out[oidx] = (sinos[idx] - val[j] * d.osc) * lw[oidx];
        }
    }
}

// ------------------------------------------------------------------ BP ---
// 64 columns x (4 * PY) rows per workgroup.  mode OUT_SET writes
// osc * A^T r; OUT_UPDATE writes clamp(x + pw * osc * A^T r).
template <int PY, bool PR = false>
__global__ __launch_bounds__(256) void batch_bp(BatchArgs B, const float *__restrict__ psino,
                                                const float *__restrict__ pw,
                                                float *__restrict__ vols, int mode, float lo,
                                                float hi) {
    const unsigned item = blockIdx.x;
    const unsigned s = find_slice(B.prefix, B.n, item);
    const SliceDesc d = B.desc[s];
    const unsigned local = item - B.prefix[s];
    const unsigned xblocks = (d.w + 63) / 64;
    const int X = (local % xblocks) * 64 + (threadIdx.x & 63);
    const int Y0 = ((local / xblocks) * 4 + (threadIdx.x >> 6)) * PY;
    if (X >= (int)d.w || Y0 >= (int)d.h) return;
    float acc[PY];
    const float *sp = psino + d.psino_off;
    const BP2DAngle *ang = B.bpa + d.ang_off;
    if (d.kind == ASTRA_AMD_GEOM_FAN)
        bp2d_column<1, PY, PR>(X, Y0, d.w, d.h, d.na, d.nd, sp, ang, acc);
    else
        bp2d_column<0, PY, PR>(X, Y0, d.w, d.h, d.na, d.nd, sp, ang, acc);
    for (int i = 0; i < PY; ++i) {
        const int Y = Y0 + i;
        if (Y >= (int)d.h) break;
        const size_t idx = d.vol_off + (size_t)Y * d.w + X;
        store_bp(vols, idx, acc[i] * d.osc, mode, pw, lo, hi);
    }
}

// The same with the sinogram footprints staged in LDS (bp2d_tile_lds).
template <int PY>
__global__ __launch_bounds__(256) void batch_bp_lds(BatchArgs B, const float *__restrict__ psino,
                                                    const float *__restrict__ pw,
                                                    float *__restrict__ vols, int mode, float lo,
                                                    float hi) {
    __shared__ Bp2Lds L;
    const unsigned item = blockIdx.x;
    const unsigned s = find_slice(B.prefix, B.n, item);
    const SliceDesc d = B.desc[s];
    const unsigned local = item - B.prefix[s];
    const unsigned xblocks = (d.w + 63) / 64;
    const int X0 = (local % xblocks) * 64;
    const int Yt0 = (local / xblocks) * 4 * PY;
    float acc[PY];
    const float *sp = psino + d.psino_off;
    const BP2DAngle *ang = B.bpa + d.ang_off;
    if (d.kind == ASTRA_AMD_GEOM_FAN)
        bp2d_tile_lds<1, PY>(X0, Yt0, d.w, d.h, d.na, d.nd, sp, ang, acc, L);
    else
        bp2d_tile_lds<0, PY>(X0, Yt0, d.w, d.h, d.na, d.nd, sp, ang, acc, L);
    const int X = X0 + (threadIdx.x & 63);
    const int Y0 = Yt0 + (threadIdx.x >> 6) * PY;
    if (X >= (int)d.w || Y0 >= (int)d.h) return;
    for (int i = 0; i < PY; ++i) {
        const int Y = Y0 + i;
        if (Y >= (int)d.h) break;
        const size_t idx = d.vol_off + (size_t)Y * d.w + X;
        store_bp(vols, idx, acc[i] * d.osc, mode, pw, lo, hi);
    }
}

// ------------------------------------------------------------ handle -----
struct Batch2D {
    unsigned n = 0;
    int device = 0;
    std::vector<SliceDesc> host;
    size_t vol_total = 0, sino_total = 0, psino_total = 0, rows_total = 0, cols_total = 0;
    std::vector<unsigned> pre_prep, pre_pad, pre_fp[7], pre_bp[3];
    DevArray desc, fpa, bpa;
    DevArray d_prep, d_pad, d_fp[7], d_bp[3];
    // work buffers (built on first use)
    DevArray line_w, pixel_w, resid, rows, cols, ones;
    bool weights_ready = false;
    double weights_relax = std::numeric_limits<double>::quiet_NaN();
    int fp_cfg = -1, bp_cfg = -1;
    // cached graph for (iterations, sinos, vols, lo, hi, fp_cfg, bp_cfg, stream)
    hipGraphExec_t exec = nullptr;
    int g_iters = -1;
    const float *g_sinos = nullptr;
    float *g_vols = nullptr;
    float g_lo = 0, g_hi = 0;
    // graphs are captured and launched on this stream (the caller's stream
    // may be the null stream, which cannot be captured); events order it
    // after / before the caller's stream
    hipStream_t cap = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    std::mutex mu;

    ~Batch2D() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (ev_in) (void)hipEventDestroy(ev_in);
        if (ev_out) (void)hipEventDestroy(ev_out);
        if (cap) (void)hipStreamDestroy(cap);
    }
};

static const int kFPDB[3] = {64, 128, 32};  // detectors per FP workgroup
static const int kBPPY[3] = {4, 2, 8};      // rows per BP thread

Batch2D *batch2d_build(unsigned n, const int *kinds, const unsigned *ws, const unsigned *hs,
                       const unsigned *nas, const unsigned *nds, const float *vecs,
                       const float *oscs) {
    AMX_REQUIRE(n > 0, "batch2d: empty batch");
    AMX_REQUIRE(kinds && ws && hs && nas && nds && vecs && oscs, "batch2d: NULL argument");
    std::unique_ptr<Batch2D> b(new Batch2D());
    b->n = n;
    AMX_HIP(hipGetDevice(&b->device));
    b->host.resize(n);
    size_t total_angles = 0;
    for (unsigned i = 0; i < n; ++i) {
        AMX_REQUIRE(kinds[i] == ASTRA_AMD_GEOM_PAR || kinds[i] == ASTRA_AMD_GEOM_FAN,
                    "batch2d: slice %u: kind must be PAR or FAN", i);
        AMX_REQUIRE(ws[i] > 0 && hs[i] > 0 && nas[i] > 0 && nds[i] > 0,
                    "batch2d: slice %u: dimensions must be positive", i);
        AMX_REQUIRE(ws[i] <= 16000 && hs[i] <= 16000 && nds[i] <= 65536,
                    "batch2d: slice %u: dimensions too large", i);
        total_angles += nas[i];
    }
    std::vector<FP2DAngle> fp(total_angles);
    std::vector<BP2DAngle> bp(total_angles), bpf(total_angles);
    b->pre_prep.assign(n + 1, 0);
    b->pre_pad.assign(n + 1, 0);
    for (int k = 0; k < 3; ++k) {
        b->pre_fp[k].assign(n + 1, 0);
        b->pre_bp[k].assign(n + 1, 0);
    }
    for (int k = 3; k < 7; ++k) b->pre_fp[k].assign(n + 1, 0);
    size_t ang = 0;
    for (unsigned i = 0; i < n; ++i) {
        SliceDesc &d = b->host[i];
        d.kind = kinds[i];
        d.w = ws[i]; d.h = hs[i]; d.na = nas[i]; d.nd = nds[i];
        d.osc = oscs[i];
        d.vol_off = b->vol_total;
        d.sino_off = b->sino_total;
        d.psino_off = b->psino_total;
        d.rows_off = b->rows_total;
        d.cols_off = b->cols_total;
        d.ang_off = (unsigned)ang;
        d.pad_ = 0;
        b->vol_total += (size_t)d.w * d.h;
        b->sino_total += (size_t)d.na * d.nd;
        b->psino_total += (size_t)d.na * (d.nd + 2 * kPad);
        b->rows_total += (size_t)d.h * (d.w + 2 * kPad);
        b->cols_total += (size_t)d.w * (d.h + 2 * kPad);
        for (unsigned a = 0; a < d.na; ++a, ++ang)
            angle2d_params(d.kind, vecs + 6 * ang, a, fp[ang], bp[ang], bpf[ang]);
        auto acc = [&](std::vector<unsigned> &p, size_t items) {
            size_t v = (size_t)p[i] + items;
            // a launch holds at most 2^32 work-items: 2^24 workgroups of 256
            AMX_REQUIRE(v < (1ull << 24), "batch2d: batch too large (%zu workgroups in one stage)", v);
            p[i + 1] = (unsigned)v;
        };
        acc(b->pre_prep, (size_t)((d.w + 2 * kPad + 63) / 64) * ((d.h + 2 * kPad + 63) / 64));
        acc(b->pre_pad, d.na);
        for (int k = 0; k < 3; ++k) {
            const unsigned db = kFPDB[k], ab = 256 / kFPDB[k];
            acc(b->pre_fp[k], (size_t)((d.nd + db - 1) / db) * ((d.na + ab - 1) / ab));
            const unsigned rows_per_group = 4 * kBPPY[k];
            acc(b->pre_bp[k], (size_t)((d.w + 63) / 64) * ((d.h + rows_per_group - 1) / rows_per_group));
        }
        acc(b->pre_fp[3], (size_t)((d.nd + 63) / 64) * ((d.na + 15) / 16));
        for (int k = 0; k < 3; ++k) {  // pair tilings: kFPDB[k] dets x 512 / kFPDB[k] angles
            const unsigned db = kFPDB[k], ab = 512 / kFPDB[k];
            acc(b->pre_fp[4 + k], (size_t)((d.nd + db - 1) / db) * ((d.na + ab - 1) / ab));
        }
    }
    b->desc = DevArray(sizeof(SliceDesc) * n);
    b->desc.upload(b->host.data(), sizeof(SliceDesc) * n);
    b->fpa = DevArray(sizeof(FP2DAngle) * total_angles);
    b->fpa.upload(fp.data(), sizeof(FP2DAngle) * total_angles);
    b->bpa = DevArray(sizeof(BP2DAngle) * total_angles);
    b->bpa.upload(bp.data(), sizeof(BP2DAngle) * total_angles);
    auto up = [&](DevArray &d, const std::vector<unsigned> &p) {
        d = DevArray(sizeof(unsigned) * p.size());
        d.upload(p.data(), sizeof(unsigned) * p.size());
    };
    up(b->d_prep, b->pre_prep);
    up(b->d_pad, b->pre_pad);
    for (int k = 0; k < 3; ++k) {
        up(b->d_fp[k], b->pre_fp[k]);
        up(b->d_bp[k], b->pre_bp[k]);
    }
    for (int k = 3; k < 7; ++k) up(b->d_fp[k], b->pre_fp[k]);
    return b.release();
}

void batch2d_destroy(Batch2D *b) { delete b; }

void batch2d_offsets(const Batch2D &b, size_t *vol_off, size_t *sino_off) {
    for (unsigned i = 0; i < b.n; ++i) {
        if (vol_off) vol_off[i] = b.host[i].vol_off;
        if (sino_off) sino_off[i] = b.host[i].sino_off;
    }
    if (vol_off) vol_off[b.n] = b.vol_total;
    if (sino_off) sino_off[b.n] = b.sino_total;
}

namespace {

BatchArgs args(const Batch2D &b, const DevArray &prefix) {
    BatchArgs a;
    a.desc = b.desc.as<SliceDesc>();
    a.n = b.n;
    a.prefix = prefix.as<unsigned>();
    a.fpa = b.fpa.as<FP2DAngle>();
    a.bpa = b.bpa.as<BP2DAngle>();
    return a;
}

void ensure_buffers(Batch2D &b) {
    if (b.resid.bytes()) return;
    b.line_w = DevArray(sizeof(float) * b.psino_total);
    b.resid = DevArray(sizeof(float) * b.psino_total);
    b.pixel_w = DevArray(sizeof(float) * b.vol_total);
    b.rows = DevArray(sizeof(float) * b.rows_total);
    b.cols = DevArray(sizeof(float) * b.cols_total);
    b.ones = DevArray(sizeof(float) * std::max(b.vol_total, b.sino_total));
    // padded sinogram buffers are written in their interior only
    AMX_HIP(hipMemset(b.line_w.as<float>(), 0, sizeof(float) * b.psino_total));
    AMX_HIP(hipMemset(b.resid.as<float>(), 0, sizeof(float) * b.psino_total));
}

void launch_prep(const Batch2D &b, const float *vols, hipStream_t s) {
    batch_prep<<<b.pre_prep[b.n], 256, 0, s>>>(args(b, b.d_prep), vols, b.rows.as<float>(),
                                                b.cols.as<float>());
    AMX_LAUNCHED();
}

void launch_fp(const Batch2D &b, int cfg, const float *sinos, float *out, int mode, hipStream_t s) {
    BatchArgs a = args(b, b.d_fp[cfg]);
    const unsigned g = b.pre_fp[cfg][b.n];
    const float *rows = b.rows.as<float>(), *cols = b.cols.as<float>();
    const float *lw = b.line_w.as<float>();
    if (cfg == 0)
        batch_fp<64><<<g, 256, 0, s>>>(a, rows, cols, sinos, lw, out, mode);
    else if (cfg == 1)
        batch_fp<128><<<g, 256, 0, s>>>(a, rows, cols, sinos, lw, out, mode);
    else if (cfg == 2)
        batch_fp<32><<<g, 256, 0, s>>>(a, rows, cols, sinos, lw, out, mode);
    else if (cfg == 3)
        batch_fp_lds<<<g, 256, 0, s>>>(a, rows, cols, sinos, lw, out, mode);
    else if (cfg == 4)
        batch_fp_pair<64><<<g, 256, 0, s>>>(a, rows, cols, sinos, lw, out, mode);
    else if (cfg == 5)
        batch_fp_pair<128><<<g, 256, 0, s>>>(a, rows, cols, sinos, lw, out, mode);
    else
        batch_fp_pair<32><<<g, 256, 0, s>>>(a, rows, cols, sinos, lw, out, mode);
    AMX_LAUNCHED();
}

// FP candidates: kFPDB tilings (0..2), the LDS-staged one (3) and the
// kFPDB tilings with two detectors per thread (4..6).  The autotuner times
// the first three; the others never won on the 4096-slice workload
// (LDS: no faster; pairs: 160 vs 173 slices/s) and run by request only.
constexpr int kFPCandidates = 7;
constexpr int kFPTimed = 3;

// BP candidates: cfg % 3 picks the rows per thread (kBPPY); cfg 0..2 read
// every tap from global memory, 3..5 from LDS-staged footprints, 6..8 from
// global memory two rows per load.  The autotuner times 0..5 (the row
// pairs never won: 175 vs 173 slices/s at best).
constexpr int kBPCandidates = 9;
constexpr int kBPTimed = 6;

void launch_bp(const Batch2D &b, int cfg, const float *psino, float *vols, int mode, float lo,
               float hi, hipStream_t s) {
    const int t = cfg % 3;
    BatchArgs a = args(b, b.d_bp[t]);
    const unsigned g = b.pre_bp[t][b.n];
    const float *pw = b.pixel_w.as<float>();
    if (cfg == 0)
        batch_bp<4><<<g, 256, 0, s>>>(a, psino, pw, vols, mode, lo, hi);
    else if (cfg == 1)
        batch_bp<2><<<g, 256, 0, s>>>(a, psino, pw, vols, mode, lo, hi);
    else if (cfg == 2)
        batch_bp<8><<<g, 256, 0, s>>>(a, psino, pw, vols, mode, lo, hi);
    else if (cfg == 3)
        batch_bp_lds<4><<<g, 256, 0, s>>>(a, psino, pw, vols, mode, lo, hi);
    else if (cfg == 4)
        batch_bp_lds<2><<<g, 256, 0, s>>>(a, psino, pw, vols, mode, lo, hi);
    else if (cfg == 5)
        batch_bp_lds<8><<<g, 256, 0, s>>>(a, psino, pw, vols, mode, lo, hi);
    else if (cfg == 6)
        batch_bp<4, true><<<g, 256, 0, s>>>(a, psino, pw, vols, mode, lo, hi);
    else if (cfg == 7)
        batch_bp<2, true><<<g, 256, 0, s>>>(a, psino, pw, vols, mode, lo, hi);
    else
        batch_bp<8, true><<<g, 256, 0, s>>>(a, psino, pw, vols, mode, lo, hi);
    AMX_LAUNCHED();
}

// SIRT weights (src/CudaSirtAlgorithm.cpp:187-222) for every slice
void precompute(Batch2D &b, double relaxation, hipStream_t s) {
    float *ones = b.ones.as<float>();
    fill_launch(ones, b.vol_total, 1.f, s);
    launch_prep(b, ones, s);
    launch_fp(b, 0, nullptr, b.line_w.as<float>(), OUT_SET, s);
    invert_launch(b.line_w.as<float>(), b.psino_total, 1.f, s);
    fill_launch(ones, b.sino_total, 1.f, s);
    batch_pad<<<b.pre_pad[b.n], 256, 0, s>>>(args(b, b.d_pad), ones, b.resid.as<float>());
    AMX_LAUNCHED();
    launch_bp(b, 0, b.resid.as<float>(), b.pixel_w.as<float>(), OUT_SET, 0.f, 0.f, s);
    invert_launch(b.pixel_w.as<float>(), b.vol_total, (float)relaxation, s);
}

// time each FP / BP tile candidate on this batch and keep the fastest
void autotune(Batch2D &b, hipStream_t s) {
    hipEvent_t e0, e1;
    AMX_HIP(hipEventCreate(&e0));
    AMX_HIP(hipEventCreate(&e1));
    float best_fp = 1e30f, best_bp = 1e30f;
    float *scratch = b.resid.as<float>();
    launch_prep(b, b.ones.as<float>(), s);
    for (int k = 0; k < kFPTimed; ++k) {
        float ms = 0.f;
        launch_fp(b, k, nullptr, scratch, OUT_SET, s);  // warm
        AMX_HIP(hipEventRecord(e0, s));
        launch_fp(b, k, nullptr, scratch, OUT_SET, s);
        AMX_HIP(hipEventRecord(e1, s));
        AMX_HIP(hipEventSynchronize(e1));
        AMX_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best_fp) {
            best_fp = ms;
            b.fp_cfg = k;
        }
    }
    float *vtmp = b.ones.as<float>();
    for (int k = 0; k < kBPTimed; ++k) {
        float ms = 0.f;
        launch_bp(b, k, scratch, vtmp, OUT_SET, 0.f, 0.f, s);
        AMX_HIP(hipEventRecord(e0, s));
        launch_bp(b, k, scratch, vtmp, OUT_SET, 0.f, 0.f, s);
        AMX_HIP(hipEventRecord(e1, s));
        AMX_HIP(hipEventSynchronize(e1));
        AMX_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best_bp) {
            best_bp = ms;
            b.bp_cfg = k;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    // ASTRA_AMD_BATCH_BP_CFG=<0..8> / ASTRA_AMD_BATCH_FP_CFG=<0..6> replace
    // the timed choice (read when a batch is first run)
    const char *eb = std::getenv("ASTRA_AMD_BATCH_BP_CFG");
    if (eb && *eb) b.bp_cfg = std::min(std::max(std::atoi(eb), 0), kBPCandidates - 1);
    const char *ef = std::getenv("ASTRA_AMD_BATCH_FP_CFG");
    if (ef && *ef) b.fp_cfg = std::min(std::max(std::atoi(ef), 0), kFPCandidates - 1);
}

void iterate(const Batch2D &b, const float *sinos, float *vols, int iters, float lo, float hi,
             hipStream_t s) {
    for (int it = 0; it < iters; ++it) {
        launch_prep(b, vols, s);
        launch_fp(b, b.fp_cfg, sinos, b.resid.as<float>(), OUT_RESID, s);
        launch_bp(b, b.bp_cfg, b.resid.as<float>(), vols, OUT_UPDATE, lo, hi, s);
    }
}

}  // namespace

void batch2d_sirt(Batch2D &b, const float *sinos, float *vols, int iters, double relaxation,
                  float lo, float hi, bool use_graph, hipStream_t s) {
    std::lock_guard<std::mutex> lk(b.mu);
    AMX_REQUIRE(sinos && vols, "batch2d_sirt: NULL data pointer");
    AMX_REQUIRE(iters >= 0, "batch2d_sirt: negative iteration count");
    ensure_buffers(b);
    if (!b.weights_ready || b.weights_relax != relaxation) {
        precompute(b, relaxation, s);
        b.weights_ready = true;
        b.weights_relax = relaxation;
    }
    if (b.fp_cfg < 0) autotune(b, s);
    if (iters == 0) return;
    if (!use_graph) {
        iterate(b, sinos, vols, iters, lo, hi, s);
        return;
    }
    if (!b.cap) {
        AMX_HIP(hipStreamCreateWithFlags(&b.cap, hipStreamNonBlocking));
        AMX_HIP(hipEventCreateWithFlags(&b.ev_in, hipEventDisableTiming));
        AMX_HIP(hipEventCreateWithFlags(&b.ev_out, hipEventDisableTiming));
    }
    // everything enqueued on the caller's stream so far precedes the graph
    AMX_HIP(hipEventRecord(b.ev_in, s));
    AMX_HIP(hipStreamWaitEvent(b.cap, b.ev_in, 0));
    const bool same = b.exec && b.g_iters == iters && b.g_sinos == sinos && b.g_vols == vols &&
                      b.g_lo == lo && b.g_hi == hi;
    if (!same) {
        if (b.exec) {
            (void)hipGraphExecDestroy(b.exec);
            b.exec = nullptr;
        }
        hipGraph_t graph = nullptr;
        AMX_HIP(hipStreamBeginCapture(b.cap, hipStreamCaptureModeThreadLocal));
        try {
            iterate(b, sinos, vols, iters, lo, hi, b.cap);
        } catch (...) {
            (void)hipStreamEndCapture(b.cap, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            throw;
        }
        AMX_HIP(hipStreamEndCapture(b.cap, &graph));
        hipError_t e = hipGraphInstantiate(&b.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            b.exec = nullptr;
            fail(ASTRA_AMD_ERR_HIP, "batch2d_sirt: graph instantiation failed: %s",
                 hipGetErrorString(e));
        }
        b.g_iters = iters;
        b.g_sinos = sinos;
        b.g_vols = vols;
        b.g_lo = lo;
        b.g_hi = hi;
    }
    ProfToken pt = prof_begin("batch2d_sirt", (double)iters, b.cap);
    AMX_HIP(hipGraphLaunch(b.exec, b.cap));
    prof_end(pt, b.cap);
    // the caller's stream continues after the graph
    AMX_HIP(hipEventRecord(b.ev_out, b.cap));
    AMX_HIP(hipStreamWaitEvent(s, b.ev_out, 0));
}

void batch2d_fp(Batch2D &b, const float *vols, float *sinos, hipStream_t s) {
    std::lock_guard<std::mutex> lk(b.mu);
    AMX_REQUIRE(sinos && vols, "batch2d_fp: NULL data pointer");
    ensure_buffers(b);
    launch_prep(b, vols, s);
    launch_fp(b, 0, nullptr, b.resid.as<float>(), OUT_SET, s);
    // padded -> packed user layout
    batch_unpad<<<b.pre_pad[b.n], 256, 0, s>>>(args(b, b.d_pad), b.resid.as<float>(), sinos);
    AMX_LAUNCHED();
}

}  // namespace amx
