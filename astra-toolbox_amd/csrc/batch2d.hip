// Batched 2D SIRT: thousands of 2D slices of uneven size, angle count and
// geometry (parallel / fan) reconstructed together.
//
// All slices live in packed device buffers.  Each pipeline stage (volume
// slice copies, residual forward projection, weighted back projection with
// the SIRT update) is ONE launch over the whole batch: a flattened table of
// work items (per-slice prefix sums) maps every workgroup to a (slice, tile)
// pair, so a 128^2 slice and a 1024^2 slice share the GPU in the same
// launch.  N iterations are captured once into a hipGraph (3N kernel nodes)
// and replayed as a single graph launch -- or run as one persistent launch
// over device work queues (batch_sirt_persistent).  FP and BP tile shapes
// are chosen per batch by timing the candidates on the batch itself.
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
// (8 waves per SIMD: the register allocation fits 64 VGPRs without spills)
template <int DB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void batch_fp(
    BatchArgs B, const float *__restrict__ rows,
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
    store_bp_rows<PY>(vols, d.vol_off + (size_t)Y0 * d.w + X, d.w, min(PY, (int)d.h - Y0), acc, d.osc,
                      mode, pw, lo, hi);
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
    store_bp_rows<PY>(vols, d.vol_off + (size_t)Y0 * d.w + X, d.w, min(PY, (int)d.h - Y0), acc, d.osc,
                      mode, pw, lo, hi);
}

// ---------------------------------------------------------- persistent ---
// The whole SIRT run as ONE launch: a grid of workgroups takes work items
// from device queues until every iteration of every slice is done.  An
// item is one unit of one stage of one slice in one iteration:
//   stage 0 (copy): one 64-row band of the padded slice copies (every
//                   64 x 64 tile of the band),
//   stage 1 (FP):   one block of DB detectors of the residual FP (every
//                   angle tile: 256 / DB angles at a time),
//   stage 2 (BP):   one band of 4 PY rows of the BP with the SIRT update
//                   (every 64-column tile).
// Per slice and stage a counter records completed items.  An item waits
// until the stage before it is complete for its slice: copy(it) after
// BP(it - 1), FP(it) after copy(it), BP(it) after FP(it).  Per slice this
// is the order of the per-stage launches, and every item runs the same
// per-tile code, so the results are the same floats.
//
// Each queue holds its items in dependency order (one iteration: every
// copy item, then every FP item, then every BP item of the queue's slices;
// iteration after iteration).  An item waits only for items of its own
// queue that were dequeued before it; the workgroups that took those are
// running and finish them without waiting on later items, so the grid
// cannot deadlock, whether or not all of it is resident.  The wait is
// bounded all the same (spin_limit polls): past it the workgroup raises
// status[0], and every workgroup stops at its next dequeue or poll, so the
// kernel always drains.
//
// There are 8 queues, one per XCD (workgroup L runs on XCD L % 8 and
// starts with queue L % 8), each with a share of the slices balanced by
// work; a workgroup whose queue is exhausted takes items from the next
// ones.  Placement only affects locality: every thread publishes its
// stores with an agent-scope release fence before the item's completion
// count is raised (release), and the waiting thread acquires at agent
// scope, so results cross XCDs (and their separate L2s) correctly.
struct PersistArgs {
    const SliceDesc *desc;
    const FP2DAngle *fpa;
    const BP2DAngle *bpa;
    const uint2 *items;  // one iteration of every queue: (slice, stage << 30 | unit)
    unsigned qoff[9];    // queue x: items[qoff[x], qoff[x + 1])
    unsigned iters, n, spin_limit;
    unsigned *heads;   // [8] dequeue counters (item index over all iterations)
    unsigned *done;    // [3][n] completed items per stage and slice
    unsigned *status;  // [0] != 0: a wait exceeded spin_limit
    const float *sinos, *lw, *pw;
    float *vols, *resid, *rows, *cols;
    float lo, hi;
};

// units of one stage of a slice with DB detectors per FP tile and PY rows
// per BP thread (host and device)
__host__ __device__ __forceinline__ unsigned persist_units(unsigned h, unsigned nd, unsigned stage,
                                                           unsigned db, unsigned py) {
    if (stage == 0) return (h + 2 * kPad + 63) / 64;
    if (stage == 1) return (nd + db - 1) / db;
    return (h + 4 * py - 1) / (4 * py);
}

union PersistLds {
    float tile[64][65];
    Bp2Lds bp;
};

// WPE: the fewest waves per SIMD the register allocation must allow (1: no
// limit; 6 caps the kernel at 80 VGPRs, with a few spills outside the
// sample loops).
template <int DB, int PY, bool LDS, int WPE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE))) void
batch_sirt_persistent(PersistArgs P) {
    __shared__ PersistLds S;
    __shared__ unsigned sh[2];
    const unsigned tid = threadIdx.x;
    const unsigned home = blockIdx.x & 7u;
    unsigned q = 0;  // queues exhausted so far (thread 0)
    for (;;) {
        if (tid == 0) {
            unsigned ok = 0, s = 0, tag = 0, it = 0;
            while (q < 8) {
                if (__hip_atomic_load(P.status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                    q = 8;
                    break;
                }
                const unsigned x = (home + q) & 7u;
                const unsigned len = P.qoff[x + 1] - P.qoff[x];
                const unsigned total = len * P.iters;
                const unsigned i =
                    total ? __hip_atomic_fetch_add(P.heads + x, 1u, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT)
                          : 0u;
                if (i < total) {
                    const uint2 e = P.items[P.qoff[x] + i % len];
                    s = e.x;
                    tag = e.y;
                    it = i / len;
                    ok = 1;
                    break;
                }
                ++q;
            }
            if (ok) {
                // wait until the slice's previous stage is complete
                const SliceDesc &d = P.desc[s];
                const unsigned stage = tag >> 30, prev = stage == 0 ? 2u : stage - 1;
                const unsigned need =
                    (stage == 0 ? it : it + 1) * persist_units(d.h, d.nd, prev, DB, PY);
                const unsigned *cnt = P.done + (size_t)prev * P.n + s;
                unsigned k = 0;
                while (__hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
                    if (++k > P.spin_limit ||
                        __hip_atomic_load(P.status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                        __hip_atomic_fetch_or(P.status, 1u, __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT);
                        ok = 0;
                        q = 8;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(4);
                }
                // the producers' stores are visible from here on
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            }
            sh[0] = ok ? s : 0xffffffffu;
            sh[1] = tag;
        }
        __syncthreads();
        // wave-uniform: the slice descriptor and its offsets go to scalar registers
        const unsigned s = __builtin_amdgcn_readfirstlane(sh[0]);
        const unsigned tag = __builtin_amdgcn_readfirstlane(sh[1]);
        if (s == 0xffffffffu) break;
        const unsigned stage = tag >> 30, unit = tag & 0x3fffffffu;
        const SliceDesc d = P.desc[s];
        if (stage == 0) {
            const unsigned tx = (d.w + 2 * kPad + 63) / 64;
#pragma unroll 1
            for (unsigned c = 0; c < tx; ++c) {
                if (c) __syncthreads();  // the previous tile's LDS is read
                vol2d_copy_tile(P.vols + d.vol_off, P.rows + d.rows_off, P.cols + d.cols_off, d.w,
                                d.h, c * 64, unit * 64, S.tile);
            }
        } else if (stage == 1) {
            constexpr int AB = 256 / DB;
            const int det = unit * DB + (tid % DB);
            const float *r = P.rows + d.rows_off, *c = P.cols + d.cols_off;
#pragma unroll 1
            for (unsigned a0 = 0; a0 < d.na; a0 += AB) {
                const int angle = a0 + (tid / DB);
                if (det >= (int)d.nd || angle >= (int)d.na) continue;
                const FP2DAngle &G = P.fpa[d.ang_off + angle];
                const float v = (d.kind == ASTRA_AMD_GEOM_FAN)
                                    ? fp2d_value<true>(G, det, 1, d.w, d.h, r, c)
                                    : fp2d_value<false>(G, det, 1, d.w, d.h, r, c);
                const size_t oidx = d.psino_off + (size_t)angle * (d.nd + 2 * kPad) + det + kPad;
                const size_t idx = d.sino_off + (size_t)angle * d.nd + det;
                P.resid[oidx] = (P.sinos[idx] - v * d.osc) * P.lw[oidx];
            }
        } else {
            const unsigned xblocks = (d.w + 63) / 64;
            const int Yt0 = unit * 4 * PY;
            const float *sp = P.resid + d.psino_off;
            const BP2DAngle *ang = P.bpa + d.ang_off;
#pragma unroll 1
            for (unsigned b = 0; b < xblocks; ++b) {
                const int X0 = b * 64;
                const int X = X0 + (tid & 63);
                const int Y0 = Yt0 + (tid >> 6) * PY;
                float acc[PY];
                if (LDS) {
                    if (d.kind == ASTRA_AMD_GEOM_FAN)
                        bp2d_tile_lds<1, PY>(X0, Yt0, d.w, d.h, d.na, d.nd, sp, ang, acc, S.bp);
                    else
                        bp2d_tile_lds<0, PY>(X0, Yt0, d.w, d.h, d.na, d.nd, sp, ang, acc, S.bp);
                }
                if (X >= (int)d.w || Y0 >= (int)d.h) continue;
                if (!LDS) {
                    if (d.kind == ASTRA_AMD_GEOM_FAN)
                        bp2d_column<1, PY>(X, Y0, d.w, d.h, d.na, d.nd, sp, ang, acc);
                    else
                        bp2d_column<0, PY>(X, Y0, d.w, d.h, d.na, d.nd, sp, ang, acc);
                }
                store_bp_rows<PY>(P.vols, d.vol_off + (size_t)Y0 * d.w + X, d.w,
                                  min(PY, (int)d.h - Y0), acc, d.osc, OUT_UPDATE, P.pw, P.lo,
                                  P.hi);
            }
        }
        // publish this thread's stores device-wide, then count the item
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __syncthreads();  // also: every thread has read sh
        if (tid == 0)
            __hip_atomic_fetch_add(P.done + (size_t)stage * P.n + s, 1u, __ATOMIC_RELEASE,
                                   __HIP_MEMORY_SCOPE_AGENT);
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
    int g_mode = -1, g_wpe = 0, g_chunk = 0;
    // persistent run: the queues (built for the tilings p_fp / p_bp), the
    // sync words (heads[8], done[3][n], status) and a pinned status copy
    DevArray p_items, p_sync;
    unsigned p_qoff[9] = {};
    int p_fp = -1, p_bp = -1, p_wpe = 1;
    unsigned p_grid = 0;
    unsigned *p_status_host = nullptr;
    std::mutex mu;

    ~Batch2D() {
        if (p_status_host) (void)hipHostFree(p_status_host);
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
// the first three; the others never won on the 4096-slice workload (LDS:
// no faster; pairs: 197 vs 200 slices/s; workgroups of both kinds in one
// launch, to keep the texture path and the VALU busy at once: 185) and run
// by request only.
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

// Queues of the persistent run for the chosen tilings: slices dealt to the
// 8 queues by decreasing work, each to the least loaded queue (work = FP
// samples + BP updates per iteration); per queue one iteration's items,
// stage after stage, slices in that order.
void build_persistent(Batch2D &b) {
    // the persistent kernel's tilings: the chosen ones, or for the
    // candidates it has no form of (LDS FP, pair and mixed FP, row-pair BP)
    // the plain tiling of the same detectors / rows per thread (the same
    // floats in every case)
    static const int fp_plain[kFPCandidates] = {0, 1, 2, 0, 0, 1, 2};
    const int pf = fp_plain[b.fp_cfg], pb = b.bp_cfg < 6 ? b.bp_cfg : b.bp_cfg - 6;
    if (b.p_items.bytes() && b.p_fp == pf && b.p_bp == pb) return;
    const unsigned db = kFPDB[pf], py = kBPPY[pb % 3];
    std::vector<unsigned> order(b.n);
    std::vector<double> work(b.n);
    for (unsigned i = 0; i < b.n; ++i) {
        const SliceDesc &d = b.host[i];
        order[i] = i;
        work[i] = (double)d.na * ((double)d.nd * std::max(d.w, d.h) + (double)d.w * d.h);
    }
    std::stable_sort(order.begin(), order.end(),
                     [&](unsigned x, unsigned y) { return work[x] > work[y]; });
    std::vector<std::vector<unsigned>> qs(8);
    double load[8] = {};
    for (unsigned i : order) {
        const int x = (int)(std::min_element(load, load + 8) - load);
        qs[x].push_back(i);
        load[x] += work[i];
    }
    std::vector<uint2> items;
    for (int x = 0; x < 8; ++x) {
        b.p_qoff[x] = (unsigned)items.size();
        for (unsigned stage = 0; stage < 3; ++stage)
            for (unsigned i : qs[x]) {
                const SliceDesc &d = b.host[i];
                const unsigned nu = persist_units(d.h, d.nd, stage, db, py);
                for (unsigned u = 0; u < nu; ++u) items.push_back(make_uint2(i, (stage << 30) | u));
            }
        AMX_REQUIRE(items.size() < (1ull << 31), "batch2d: too many persistent work items");
    }
    b.p_qoff[8] = (unsigned)items.size();
    b.p_items = DevArray(sizeof(uint2) * items.size());
    b.p_items.upload(items.data(), sizeof(uint2) * items.size());
    if (!b.p_sync.bytes()) b.p_sync = DevArray(sizeof(unsigned) * (8 + 3 * (size_t)b.n + 1));
    if (!b.p_status_host)
        AMX_HIP(hipHostMalloc((void **)&b.p_status_host, sizeof(unsigned), hipHostMallocDefault));
    b.p_fp = pf;
    b.p_bp = pb;
}

// Most polls of a dependency wait (about 2 us each) before a persistent run
// gives up; ASTRA_AMD_BATCH_SPIN_LIMIT overrides it (read at every run).
unsigned persist_spin_limit() {
    const char *e = std::getenv("ASTRA_AMD_BATCH_SPIN_LIMIT");
    if (e && *e) return (unsigned)std::min<unsigned long>(std::strtoul(e, nullptr, 10), 0xffffffffu);
    return 1u << 23;
}

// Waves per SIMD asked of the persistent kernel's register allocation
// (ASTRA_AMD_BATCH_PERSIST_WPE: 6, or 1 for no limit; read at every run).
int persist_wpe() {
    const char *e = std::getenv("ASTRA_AMD_BATCH_PERSIST_WPE");
    return (e && std::atoi(e) == 6) ? 6 : 1;
}

// P == NULL: size the grid of the kernel for (tilings, wpe) -- as many
// workgroups as are resident at once, a multiple of 8 -- into *grid
// (before a capture: the occupancy query is not a stream operation, but
// it is kept out of the captured region all the same).  Else launch it
// with *grid workgroups.
template <int DB, int PY, bool LDS>
void persist_t(const PersistArgs *P, int wpe, unsigned *grid, hipStream_t s) {
    auto kern = wpe == 6 ? batch_sirt_persistent<DB, PY, LDS, 6>
                         : batch_sirt_persistent<DB, PY, LDS, 1>;
    if (!P) {
        int per_cu = 0, cus = 0, dev = 0;
        AMX_HIP(hipGetDevice(&dev));
        AMX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 256, 0));
        AMX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        *grid = std::max(8u, (unsigned)(std::max(per_cu, 1) * std::max(cus, 1)) & ~7u);
        return;
    }
    kern<<<*grid, 256, 0, s>>>(*P);
    AMX_LAUNCHED();
}

void persist_dispatch(const Batch2D &b, const PersistArgs *P, int wpe, unsigned *grid,
                      hipStream_t s) {
    switch (b.p_fp * 6 + b.p_bp) {
#define AMX_PERSIST_CASE(F, B, DB, PY, LDS) \
    case F * 6 + B: persist_t<DB, PY, LDS>(P, wpe, grid, s); break;
        AMX_PERSIST_CASE(0, 0, 64, 4, false)
        AMX_PERSIST_CASE(0, 1, 64, 2, false)
        AMX_PERSIST_CASE(0, 2, 64, 8, false)
        AMX_PERSIST_CASE(0, 3, 64, 4, true)
        AMX_PERSIST_CASE(0, 4, 64, 2, true)
        AMX_PERSIST_CASE(0, 5, 64, 8, true)
        AMX_PERSIST_CASE(1, 0, 128, 4, false)
        AMX_PERSIST_CASE(1, 1, 128, 2, false)
        AMX_PERSIST_CASE(1, 2, 128, 8, false)
        AMX_PERSIST_CASE(1, 3, 128, 4, true)
        AMX_PERSIST_CASE(1, 4, 128, 2, true)
        AMX_PERSIST_CASE(1, 5, 128, 8, true)
        AMX_PERSIST_CASE(2, 0, 32, 4, false)
        AMX_PERSIST_CASE(2, 1, 32, 2, false)
        AMX_PERSIST_CASE(2, 2, 32, 8, false)
        AMX_PERSIST_CASE(2, 3, 32, 4, true)
        AMX_PERSIST_CASE(2, 4, 32, 2, true)
        AMX_PERSIST_CASE(2, 5, 32, 8, true)
#undef AMX_PERSIST_CASE
        default: fail(ASTRA_AMD_ERR_INTERNAL, "batch2d_sirt: no persistent kernel for the tilings");
    }
}

// Iterations per persistent launch: as many as keep every queue's dequeue
// counter below 2^31 (ASTRA_AMD_BATCH_PERSIST_ITERS caps it further; read
// at every run).
int persist_chunk(const Batch2D &b) {
    size_t longest = 1;
    for (int x = 0; x < 8; ++x) longest = std::max<size_t>(longest, b.p_qoff[x + 1] - b.p_qoff[x]);
    long chunk = (long)(((1ull << 31) - 1 - (1u << 24)) / longest);
    const char *e = std::getenv("ASTRA_AMD_BATCH_PERSIST_ITERS");
    if (e && std::atol(e) > 0) chunk = std::min(chunk, std::atol(e));
    AMX_REQUIRE(chunk >= 1, "batch2d_sirt: too many work items for a persistent launch");
    return (int)std::min<long>(chunk, 1 << 30);
}

// One persistent run: clear the sync words (a fill kernel, so the run can
// be captured), then the persistent kernel (b.p_wpe, b.p_grid) -- one
// launch, or consecutive launches of `chunk` iterations each when the
// dequeue counters would overflow (the stop flag is cleared only before
// the first, so a stopped launch stops the ones after it).
void iterate_persistent(const Batch2D &b, const float *sinos, float *vols, int iters, int chunk,
                        float lo, float hi, hipStream_t s) {
    unsigned *sync = b.p_sync.as<unsigned>();
    PersistArgs P;
    P.desc = b.desc.as<SliceDesc>();
    P.fpa = b.fpa.as<FP2DAngle>();
    P.bpa = b.bpa.as<BP2DAngle>();
    P.items = b.p_items.as<uint2>();
    for (int x = 0; x <= 8; ++x) P.qoff[x] = b.p_qoff[x];
    P.n = b.n;
    P.spin_limit = persist_spin_limit();
    P.heads = sync;
    P.done = sync + 8;
    P.status = sync + 8 + 3 * (size_t)b.n;
    P.sinos = sinos;
    P.lw = b.line_w.as<float>();
    P.pw = b.pixel_w.as<float>();
    P.vols = vols;
    P.resid = b.resid.as<float>();
    P.rows = b.rows.as<float>();
    P.cols = b.cols.as<float>();
    P.lo = lo;
    P.hi = hi;
    for (int done = 0; done < iters; done += chunk) {
        const size_t words = 8 + 3 * (size_t)b.n + (done == 0 ? 1 : 0);
        fill_launch(reinterpret_cast<float *>(sync), words, 0.f, s);
        P.iters = (unsigned)std::min(chunk, iters - done);
        unsigned grid = b.p_grid;
        persist_dispatch(b, &P, b.p_wpe, &grid, s);
    }
}

}  // namespace

void batch2d_sirt(Batch2D &b, const float *sinos, float *vols, int iters, double relaxation,
                  float lo, float hi, int run_mode, hipStream_t s) {
    std::lock_guard<std::mutex> lk(b.mu);
    AMX_REQUIRE(sinos && vols, "batch2d_sirt: NULL data pointer");
    AMX_REQUIRE(iters >= 0, "batch2d_sirt: negative iteration count");
    AMX_REQUIRE(run_mode >= 0 && run_mode <= 2, "batch2d_sirt: run mode must be 0, 1 or 2");
    ensure_buffers(b);
    if (!b.weights_ready || b.weights_relax != relaxation) {
        precompute(b, relaxation, s);
        b.weights_ready = true;
        b.weights_relax = relaxation;
    }
    if (b.fp_cfg < 0) autotune(b, s);
    if (iters == 0) return;
    const bool persistent = run_mode == 2;
    const int wpe = persistent ? persist_wpe() : 1;
    int chunk = 0;
    if (persistent) {
        build_persistent(b);
        persist_dispatch(b, nullptr, wpe, &b.p_grid, s);
        b.p_wpe = wpe;
        chunk = persist_chunk(b);
    }
    if (run_mode == 0) {
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
                      b.g_lo == lo && b.g_hi == hi && b.g_mode == run_mode && b.g_wpe == wpe &&
                      b.g_chunk == chunk;
    if (!same) {
        if (b.exec) {
            (void)hipGraphExecDestroy(b.exec);
            b.exec = nullptr;
        }
        hipGraph_t graph = nullptr;
        AMX_HIP(hipStreamBeginCapture(b.cap, hipStreamCaptureModeThreadLocal));
        try {
            if (persistent)
                iterate_persistent(b, sinos, vols, iters, chunk, lo, hi, b.cap);
            else
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
        b.g_mode = run_mode;
        b.g_wpe = wpe;
        b.g_chunk = chunk;
    }
    ProfToken pt = prof_begin(persistent ? "batch2d_sirt_persistent" : "batch2d_sirt",
                              (double)iters, b.cap);
    AMX_HIP(hipGraphLaunch(b.exec, b.cap));
    prof_end(pt, b.cap);
    if (persistent) {
        // the run ends before the call returns; a wait past its limit
        // stopped it (the sync words are cleared by the next run)
        AMX_HIP(hipMemcpyAsync(b.p_status_host, b.p_sync.as<unsigned>() + 8 + 3 * (size_t)b.n,
                               sizeof(unsigned), hipMemcpyDeviceToHost, b.cap));
        AMX_HIP(hipStreamSynchronize(b.cap));
        AMX_REQUIRE(*b.p_status_host == 0,
                    "batch2d_sirt: persistent run stopped: a work item waited past the spin "
                    "limit (%u polls); the volumes hold a partial result",
                    persist_spin_limit());
    }
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
