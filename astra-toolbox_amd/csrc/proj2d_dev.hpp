// Device code shared by the single-slice 2D projectors (proj2d.hip) and the
// batched 2D SIRT (batch2d.hip): per-ray forward projection, per-pixel back
// projection, and the output epilogues.
//
// Sampling semantics follow cuda/2d/par_fp.cu:53-200, cuda/2d/fan_fp.cu:55-193,
// cuda/2d/par_bp.cu:52-87 and cuda/2d/fan_bp.cu:57-103.
#pragma once

#include <hip/hip_runtime.h>

#include "devutil.hpp"
#include "internal.hpp"

namespace amx {

struct Ray2D {
    float a, c, scale;
    int lo, hi;
};

// Per-ray orientation.  Parallel: per angle (cuda/2d/par_fp.cu:272).  Fan:
// per ray in float32, adding 0.5 in world units as the reference does
// (cuda/2d/fan_fp.cu:79-83); DetS + det * U is one fused multiply-add.
template <bool FAN>
__device__ __forceinline__ bool ray_vertical(const FP2DAngle &G, int det) {
    if (!FAN) return G.vertical != 0;
    const float sx = G.raw[0], sy = G.raw[1], dsx = G.raw[2], dsy = G.raw[3];
    const float ux = G.raw[4], uy = G.raw[5];
    float tx = fmaf((float)det, ux, dsx);
    tx = tx + 0.5f;
    tx = tx - sx;
    float ty = fmaf((float)det, uy, dsy);
    ty = ty + 0.5f;
    ty = ty - sy;
    return fabsf(ty) > fabsf(tx);
}

// FP64 setup of one (sub-)ray through detector position `fdet` (pixel
// units from the detector corner).  t(s) = a s + c is the padded texel
// coordinate along the slice at slice index s.
template <bool FAN>
__device__ __forceinline__ Ray2D ray2d(const FP2DAngle &G, double fdet, bool vertical,
                                      unsigned w, unsigned h) {
    const double P0 = G.p0[0] + (fdet - 0.5) * G.u[0];
    const double P1 = G.p0[1] + (fdet - 0.5) * G.u[1];
    const double rx = FAN ? G.r[0] - P0 : G.r[0];
    const double ry = FAN ? G.r[1] - P1 : G.r[1];
    Ray2D R;
    R.lo = 0;
    double a, c;
    const double len = sqrt(rx * rx + ry * ry);
    if (vertical) {
        // rows are slices (world y = -(r - h/2 + 0.5)); coordinate along columns
        const double k = rx / ry;
        a = -k;
        c = a * (-0.5 * h + 0.5) + P0 - P1 * k + 0.5 * w - 0.5;
        R.hi = (int)h - 1;
        clip_range(a, c, (double)w, (int)h, R.lo, R.hi);
        R.scale = (float)(len / fabs(ry));
    } else {
        // columns are slices (world x = s - w/2 + 0.5); coordinate along rows
        const double k = ry / rx;
        a = -k;
        c = a * (-0.5 * w + 0.5) + 0.5 * h - 0.5 - P1 + P0 * k;
        R.hi = (int)w - 1;
        clip_range(a, c, (double)h, (int)w, R.lo, R.hi);
        R.scale = (float)(len / fabs(rx));
    }
    R.a = (float)a;
    R.c = (float)(c + kPad);
    return R;
}

// Sum of linear samples along a ray over slices [lo, hi] of a padded
// slice-major copy (slice pitch `pitch`, coordinate clamp `hi_t`).
__device__ __forceinline__ float march2d(const float *__restrict__ base, unsigned pitch,
                                         float hi_t, const Ray2D &R) {
    float acc0 = 0.f, acc1 = 0.f;
    const float *p = base + (size_t)(R.lo > 0 ? R.lo : 0) * pitch;
    int s = R.lo;
    for (; s + 1 <= R.hi; s += 2) {
        const float sf = (float)s;
        float t0 = clamp_pad(fmaf(R.a, sf, R.c), hi_t);
        float t1 = clamp_pad(fmaf(R.a, sf + 1.f, R.c), hi_t);
        float f0 = floorf(t0), f1 = floorf(t1);
        float2 q0 = ld2b(p, (unsigned)f0 << 2), q1 = ld2b(p + pitch, (unsigned)f1 << 2);
        acc0 += lerp(q0.x, q0.y, t0 - f0);
        acc1 += lerp(q1.x, q1.y, t1 - f1);
        p += 2 * pitch;
    }
    if (s <= R.hi) {
        float t0 = clamp_pad(fmaf(R.a, (float)s, R.c), hi_t);
        float f0 = floorf(t0);
        float2 q0 = ld2b(p, (unsigned)f0 << 2);
        acc0 += lerp(q0.x, q0.y, t0 - f0);
    }
    return acc0 + acc1;
}

// Forward projection value of detector `det` (before the output scale):
// mean over `rays` sub-rays of (ray sum x ray-length weight).
template <bool FAN>
__device__ __forceinline__ float fp2d_value(const FP2DAngle &G, int det, unsigned rays,
                                            unsigned w, unsigned h, const float *rows,
                                            const float *cols) {
    const bool vertical = ray_vertical<FAN>(G, det);
    float sum = 0.f;
    for (unsigned k = 0; k < rays; ++k) {
        const double fdet = det + (0.5 + k) / (double)rays;
        Ray2D R = ray2d<FAN>(G, fdet, vertical, w, h);
        float v = vertical ? march2d(rows, w + 2 * kPad, (float)w + 2.5f, R)
                           : march2d(cols, h + 2 * kPad, (float)h + 2.5f, R);
        sum += v * R.scale;
    }
    return sum / (float)rays;
}

// out[oidx] per mode (OUT_ADD / OUT_SET / OUT_RESID / OUT_ADDMUL)
__device__ __forceinline__ void store_fp(float *out, size_t oidx, float val, int mode,
                                         const float *rb, size_t idx, const float *rw) {
    if (mode == OUT_ADD) {
        out[oidx] += val;
    } else if (mode == OUT_SET) {
        out[oidx] = val;
    } else if (mode == OUT_RESID) {
        float r = rb[idx] + val;
        if (rw) r *= rw[oidx];
        out[oidx] = r;
    } else if (mode == OUT_RATIO) {
        out[oidx] = em_ratio(rb[idx], val);
    } else {
        float r = out[oidx] + val;
        if (rw) r *= rw[oidx];
        out[oidx] = r;
    }
}

// Back projection of PY consecutive rows (Y0 .. Y0+PY-1) of column X over
// all angles from a padded sinogram [na][nd+4].
// KIND: 0 parallel (per-angle weight), 1 fan (fr), 2 fan FBP (fr^2),
// 3 parallel unweighted, 4 fan unweighted (SART's BP, whose per-angle and
// fr weights cancel against its line weights: cuda/2d/par_bp.cu:139-159,
// cuda/2d/fan_bp.cu:167-199).
template <int KIND, int PY>
__device__ __forceinline__ void bp2d_column(int X, int Y0, unsigned w, unsigned h, unsigned na,
                                            unsigned nd, const float *__restrict__ sp,
                                            const BP2DAngle *__restrict__ ang, float *acc) {
    const float fx = X - 0.5f * w + 0.5f;
    const float fy0 = -(Y0 - 0.5f * h + 0.5f);  // world y of row Y0
    const unsigned pitch = nd + 2 * kPad;
    const float hi = (float)nd + 2.5f;
#pragma unroll
    for (int i = 0; i < PY; ++i) acc[i] = 0.f;
    const float *row = sp;
    for (unsigned a = 0; a < na; ++a, row += pitch) {
        const BP2DAngle c = ang[a];
        float num = fmaf(fy0, c.num[2], fmaf(fx, c.num[1], c.num[0]));
        float den = fmaf(fy0, c.den[2], fmaf(fx, c.den[1], c.den[0]));
#pragma unroll
        for (int i = 0; i < PY; ++i) {
            float fr = (KIND == 0 || KIND == 3) ? 1.f : __builtin_amdgcn_rcpf(den);
            float t = clamp_pad(num * fr, hi);
            float f = floorf(t);
            float2 q = ld2b(row, (unsigned)f << 2);
            float val = lerp(q.x, q.y, t - f);
            float wgt = (KIND == 0) ? c.w : (KIND == 1 ? fr : (KIND == 2 ? fr * fr : 1.f));
            acc[i] = fmaf(wgt, val, acc[i]);
            num -= c.num[2];  // next row: world y decreases by 1
            den -= c.den[2];
        }
    }
}

__device__ __forceinline__ void store_bp(float *vol, size_t idx, float val, int mode,
                                         const float *uw, float lo, float hi) {
    if (mode == OUT_ADD) {
        vol[idx] += val;
    } else if (mode == OUT_SET) {
        vol[idx] = val;
    } else if (mode == OUT_EMUPD) {
        vol[idx] *= val * uw[idx];
    } else {  // OUT_UPDATE: clamp(x + w * val)
        float wv = uw ? uw[idx] : 1.f;
        vol[idx] = fminf(fmaxf(fmaf(wv, val, vol[idx]), lo), hi);
    }
}

// One 64 x 64 tile of the padded slice copies of a [h][w] volume: rows copy
// [h][w+4] and columns copy [w][h+4]; tile origin (c0, r0) in padded
// coordinates.  Uses a 64 x 65 LDS tile; 256 threads.
__device__ __forceinline__ void vol2d_copy_tile(const float *__restrict__ vol,
                                                float *__restrict__ rows,
                                                float *__restrict__ cols, unsigned w, unsigned h,
                                                int c0, int r0, float (*tile)[65]) {
    const int pw = w + 2 * kPad, ph = h + 2 * kPad;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int i = ty; i < 64; i += 4) {
        int pr = r0 + i, pc = c0 + tx;
        int r = pr - kPad, c = pc - kPad;
        float v = 0.f;
        if (r >= 0 && r < (int)h && c >= 0 && c < (int)w) v = vol[(size_t)r * w + c];
        tile[i][tx] = v;
        if (r >= 0 && r < (int)h && pc < pw) rows[(size_t)r * pw + pc] = v;
    }
    __syncthreads();
    for (int i = ty; i < 64; i += 4) {
        int pc = c0 + i, pr = r0 + tx;
        int c = pc - kPad;
        if (c >= 0 && c < (int)w && pr < ph) cols[(size_t)c * ph + pr] = tile[tx][i];
    }
}

}  // namespace amx
