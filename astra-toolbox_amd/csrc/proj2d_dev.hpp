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
// slice-major copy (slice pitch `pitch` texels, coordinate clamp `hi_t`),
// read through a buffer resource over the copy (texel records: one
// v_add3 of per-lane address per sample).  Samples at even offsets from lo
// go to acc0, odd ones to acc1; four loads are issued before the first of
// them is used (the loop is bound by the loads' latency).  The clamped
// coordinate t >= 0.5 truncates to its floor, and frac(t) = t - floor(t).
__device__ __forceinline__ float march2d(__amdgpu_buffer_rsrc_t rs, unsigned pitch, float hi_t,
                                         const Ray2D &R) {
    float acc0 = 0.f, acc1 = 0.f;
    unsigned row = (unsigned)(R.lo > 0 ? R.lo : 0) * pitch;  // first texel of slice s
    int s = R.lo;
    for (; s + 3 <= R.hi; s += 4) {
        float2 q[4];
        float wt[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float t = clamp_pad(fmaf(R.a, (float)(s + k), R.c), hi_t);
            q[k] = ld2i(rs, row + k * pitch + (unsigned)t);
            wt[k] = frac(t);
        }
        acc0 += lerp(q[0].x, q[0].y, wt[0]);
        acc1 += lerp(q[1].x, q[1].y, wt[1]);
        acc0 += lerp(q[2].x, q[2].y, wt[2]);
        acc1 += lerp(q[3].x, q[3].y, wt[3]);
        row += 4 * pitch;
    }
    if (s + 1 <= R.hi) {
        const float sf = (float)s;
        float t0 = clamp_pad(fmaf(R.a, sf, R.c), hi_t);
        float t1 = clamp_pad(fmaf(R.a, sf + 1.f, R.c), hi_t);
        float2 q0 = ld2i(rs, row + (unsigned)t0), q1 = ld2i(rs, row + pitch + (unsigned)t1);
        acc0 += lerp(q0.x, q0.y, frac(t0));
        acc1 += lerp(q1.x, q1.y, frac(t1));
        row += 2 * pitch;
        s += 2;
    }
    if (s <= R.hi) {
        float t0 = clamp_pad(fmaf(R.a, (float)s, R.c), hi_t);
        float2 q0 = ld2i(rs, row + (unsigned)t0);
        acc0 += lerp(q0.x, q0.y, frac(t0));
    }
    return acc0 + acc1;
}

// Sums of two rays along the same slice axis of one padded copy, over the
// union of their slice ranges: outside its own range a ray's coordinate is
// clamped into the zero border, so the extra samples add exact zeros, and
// one alternation of accumulators over the union splits each ray's
// samples into the same two sums as march2d (possibly with the two roles
// swapped), so each result equals march2d's.  One 16-byte load from the
// lower of the two left texels serves both rays when their left texels
// are at most two apart (the load starts at most pitch - 4, so it stays
// inside the row); otherwise the other ray takes its own 8-byte load.
__device__ __forceinline__ float2 march2d_pair(__amdgpu_buffer_rsrc_t rs, unsigned pitch,
                                               float hi_t, const Ray2D &A, const Ray2D &B) {
    const bool ea = A.lo > A.hi, eb = B.lo > B.hi;
    if (ea && eb) return make_float2(0.f, 0.f);
    const int lo = ea ? B.lo : (eb ? A.lo : min(A.lo, B.lo));
    const int hi = ea ? B.hi : (eb ? A.hi : max(A.hi, B.hi));
    const float gmax = (float)pitch - 4.f;
    float a0 = 0.f, a1 = 0.f, b0 = 0.f, b1 = 0.f;
    unsigned row = (unsigned)(lo > 0 ? lo : 0) * pitch;
    for (int s = lo; s <= hi; ++s, row += pitch) {
        const float sf = (float)s;
        const float ta = clamp_pad(fmaf(A.a, sf, A.c), hi_t);
        const float tb = clamp_pad(fmaf(B.a, sf, B.c), hi_t);
        const float fa = floorf(ta), fb = floorf(tb);
        const float g = fminf(fminf(fa, fb), gmax);
        const float4 x = ld4i(rs, row + (unsigned)g);
        const float da = fa - g, db = fb - g;
        float2 qa = pick2(x, da), qb = pick2(x, db);
        if (da > 2.5f) qa = ld2i(rs, row + (unsigned)fa);
        if (db > 2.5f) qb = ld2i(rs, row + (unsigned)fb);
        const float va = lerp(qa.x, qa.y, ta - fa), vb = lerp(qb.x, qb.y, tb - fb);
        if ((s - lo) & 1) {
            a1 += va;
            b1 += vb;
        } else {
            a0 += va;
            b0 += vb;
        }
    }
    return make_float2(a0 + a1, b0 + b1);
}

// Forward projection value of detector `det` (before the output scale):
// mean over `rays` sub-rays of (ray sum x ray-length weight).
template <bool FAN>
__device__ __forceinline__ float fp2d_value(const FP2DAngle &G, int det, unsigned rays,
                                            unsigned w, unsigned h, const float *rows,
                                            const float *cols) {
    // resources over the two copies, built from their (wave-uniform)
    // bases before the per-ray choice of axis
    const __amdgpu_buffer_rsrc_t rr = texel_rsrc(rows), rc = texel_rsrc(cols);
    const bool vertical = ray_vertical<FAN>(G, det);
    float sum = 0.f;
    for (unsigned k = 0; k < rays; ++k) {
        const double fdet = det + (0.5 + k) / (double)rays;
        Ray2D R = ray2d<FAN>(G, fdet, vertical, w, h);
        float v = vertical ? march2d(rr, w + 2 * kPad, (float)w + 2.5f, R)
                           : march2d(rc, h + 2 * kPad, (float)h + 2.5f, R);
        sum += v * R.scale;
    }
    return sum / (float)rays;
}

// Values of detectors det and det + 1 (one sub-ray each), the two rays
// marched together when they run along the same slice axis; equal to two
// fp2d_value calls.  det + 1 may be past the detector (its value is unused).
template <bool FAN>
__device__ __forceinline__ float2 fp2d_value_pair(const FP2DAngle &G, int det, unsigned w,
                                                  unsigned h, const float *rows,
                                                  const float *cols) {
    const bool va = ray_vertical<FAN>(G, det), vb = ray_vertical<FAN>(G, det + 1);
    if (va != vb)
        return make_float2(fp2d_value<FAN>(G, det, 1, w, h, rows, cols),
                           fp2d_value<FAN>(G, det + 1, 1, w, h, rows, cols));
    const __amdgpu_buffer_rsrc_t rr = texel_rsrc(rows), rc = texel_rsrc(cols);
    const Ray2D A = ray2d<FAN>(G, det + 0.5, va, w, h);
    const Ray2D B = ray2d<FAN>(G, det + 1.5, va, w, h);
    const float2 s = va ? march2d_pair(rr, w + 2 * kPad, (float)w + 2.5f, A, B)
                        : march2d_pair(rc, h + 2 * kPad, (float)h + 2.5f, A, B);
    // fp2d_value: (0 + sum * scale) / 1
    return make_float2(s.x * A.scale, s.y * B.scale);
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
// PR: rows taken in pairs, one 16-byte load from the lower of the two left
// texels serving both rows when those are at most two apart (the load
// starts at most pitch - 4, inside the row), else an 8-byte load for the
// far one; the same taps and values as PR = false.
template <int KIND, int PY, bool PR = false>
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
        // every tap load of the PY rows issued before the first use (the
        // loop is bound by the loads' latency)
        float2 q[PY];
        float wt[PY], wg[PY], fl[PY];
#pragma unroll
        for (int i = 0; i < PY; ++i) {
            float fr = (KIND == 0 || KIND == 3) ? 1.f : __builtin_amdgcn_rcpf(den);
            float t = clamp_pad(num * fr, hi);
            // t >= 0.5: truncation is the floor, frac(t) = t - floor(t)
            if (PR) fl[i] = floorf(t);
            if (!PR) q[i] = ld2b(row, (unsigned)t << 2);
            wt[i] = frac(t);
            wg[i] = (KIND == 0) ? c.w : (KIND == 1 ? fr : (KIND == 2 ? fr * fr : 1.f));
            num -= c.num[2];  // next row: world y decreases by 1
            den -= c.den[2];
        }
        if (PR) {
            static_assert(!PR || PY % 2 == 0, "row pairs");
            const float gmax = (float)pitch - 4.f;
#pragma unroll
            for (int i = 0; i < PY; i += 2) {
                const float g = fminf(fminf(fl[i], fl[i + 1]), gmax);
                const float4 x = ld4u(row, (unsigned)g << 2);
                const float d0 = fl[i] - g, d1 = fl[i + 1] - g;
                q[i] = pick2(x, d0);
                q[i + 1] = pick2(x, d1);
                if (d0 > 2.5f) q[i] = ld2b(row, (unsigned)fl[i] << 2);
                if (d1 > 2.5f) q[i + 1] = ld2b(row, (unsigned)fl[i + 1] << 2);
            }
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int i = 0; i < PY; ++i) acc[i] = fmaf(wg[i], lerp(q[i].x, q[i].y, wt[i]), acc[i]);
    }
}

// ---- forward projection with slice rows staged in LDS -------------------
// A workgroup of 256 threads covers 64 detectors x 16 angles: thread t takes
// detector d0 + (t & 63) at the 4 angles a0 + 4 (t >> 6) + j, j < 4.  When
// every ray of the group runs along the same axis (rows or columns), the
// group marches the slices in lockstep, kFp2SC at a time: the interval of
// padded texels that its rays sample in those slices (bounding interval of
// each ray's coordinate at the chunk's ends, reduced over the group,
// widened by one texel) is copied into LDS with coalesced loads, and the
// samples read LDS.  A chunk whose interval exceeds kFp2Wcap texels, and a
// group with rays along both axes, sample global memory.  Each ray sums its
// samples exactly as march2d does (alternating accumulators from its own
// first slice), so the values equal fp2d_value's.
constexpr int kFp2SC = 16;     // slices per staged chunk
constexpr int kFp2Wcap = 256;  // texels per staged slice row
constexpr int kFp2Rays = 4;    // angles per thread

struct Fp2Lds {
    float buf[kFp2SC * kFp2Wcap];
    int red[2][2][4];  // [chunk parity][min, max][wave]
};

template <bool FAN>
__device__ __forceinline__ void fp2d_tile_lds(int det, int a0, unsigned na, unsigned nd, unsigned w,
                                              unsigned h, const FP2DAngle *__restrict__ fpa,
                                              const float *__restrict__ rows,
                                              const float *__restrict__ cols, float (&val)[kFp2Rays],
                                              Fp2Lds &L) {
    const unsigned tid = threadIdx.x, wave = tid >> 6;
    Ray2D R[kFp2Rays];
    bool on[kFp2Rays];
    int vert = 0, horiz = 0;
#pragma unroll
    for (int j = 0; j < kFp2Rays; ++j) {
        const int angle = a0 + j;
        on[j] = det < (int)nd && angle < (int)na;
        R[j].a = 0.f;
        R[j].c = 0.5f;
        R[j].scale = 0.f;
        R[j].lo = 1;
        R[j].hi = 0;
        val[j] = 0.f;
        if (on[j]) {
            const FP2DAngle &G = fpa[angle];
            const bool v = ray_vertical<FAN>(G, det);
            R[j] = ray2d<FAN>(G, det + 0.5, v, w, h);
            vert |= v ? 1 : 0;
            horiz |= v ? 0 : 1;
        }
    }
    const __amdgpu_buffer_rsrc_t rr = texel_rsrc(rows), rc = texel_rsrc(cols);
    const bool any_v = __syncthreads_or(vert), any_h = __syncthreads_or(horiz);
    if (any_v && any_h) {
        // rays along both axes: each ray from global memory
#pragma unroll
        for (int j = 0; j < kFp2Rays; ++j) {
            if (!on[j]) continue;
            const bool v = ray_vertical<FAN>(fpa[a0 + j], det);
            const float s = v ? march2d(rr, w + 2 * kPad, (float)w + 2.5f, R[j])
                              : march2d(rc, h + 2 * kPad, (float)h + 2.5f, R[j]);
            val[j] = s * R[j].scale;
        }
        return;
    }
    if (!any_v && !any_h) return;  // no ray in the group
    const bool V = any_v;
    const float *base = V ? rows : cols;
    const unsigned pitch = V ? w + 2 * kPad : h + 2 * kPad;
    const float hi_t = V ? (float)w + 2.5f : (float)h + 2.5f;
    // slice range of the group
    int lo = 1 << 30, hi = -1;
#pragma unroll
    for (int j = 0; j < kFp2Rays; ++j)
        if (on[j] && R[j].lo <= R[j].hi) {
            lo = min(lo, R[j].lo);
            hi = max(hi, R[j].hi);
        }
    lo = wave_min(lo);
    hi = wave_max(hi);
    if ((tid & 63) == 0) {
        L.red[1][0][wave] = lo;
        L.red[1][1][wave] = hi;
    }
    __syncthreads();
    int glo = L.red[1][0][0], ghi = L.red[1][1][0];
#pragma unroll
    for (int k = 1; k < 4; ++k) {
        glo = min(glo, L.red[1][0][k]);
        ghi = max(ghi, L.red[1][1][k]);
    }
    float acc0[kFp2Rays], acc1[kFp2Rays];
#pragma unroll
    for (int j = 0; j < kFp2Rays; ++j) acc0[j] = acc1[j] = 0.f;
    int parity = 0;
    for (int cs = glo; cs <= ghi; cs += kFp2SC, parity ^= 1) {
        const int ce = min(cs + kFp2SC - 1, ghi);
        // interval of texels sampled in slices [cs, ce]
        int tlo = 1 << 30, thi = -1;
#pragma unroll
        for (int j = 0; j < kFp2Rays; ++j) {
            const int sb = max(cs, R[j].lo), se = min(ce, R[j].hi);
            if (on[j] && sb <= se) {
                const float tb = clamp_pad(fmaf(R[j].a, (float)sb, R[j].c), hi_t);
                const float te = clamp_pad(fmaf(R[j].a, (float)se, R[j].c), hi_t);
                tlo = min(tlo, (int)fminf(tb, te));
                thi = max(thi, (int)fmaxf(tb, te));
            }
        }
        tlo = wave_min(tlo);
        thi = wave_max(thi);
        // L.red[parity] was last read before the previous chunk's second
        // barrier
        if ((tid & 63) == 0) {
            L.red[parity][0][wave] = tlo;
            L.red[parity][1][wave] = thi;
        }
        __syncthreads();  // also: every thread is done with the previous chunk's rows
        int t0 = L.red[parity][0][0], t1 = L.red[parity][1][0];
#pragma unroll
        for (int k = 1; k < 4; ++k) {
            t0 = min(t0, L.red[parity][0][k]);
            t1 = max(t1, L.red[parity][1][k]);
        }
        t0 = max(0, t0 - 1);
        t1 = min((int)pitch - 1, t1 + 2);
        const int wd = t1 - t0 + 1;
        const bool staged = t1 >= t0 && wd <= kFp2Wcap;
        if (staged) {
            const int nrows = ce - cs + 1;
            for (int i = (int)tid; i < nrows * kFp2Wcap; i += 256) {
                const int r = i / kFp2Wcap, c = i & (kFp2Wcap - 1);
                if (c < wd) L.buf[i] = base[(size_t)(cs + r) * pitch + t0 + c];
            }
        }
        __syncthreads();
        for (int sl = cs; sl <= ce; ++sl) {
            const float sf = (float)sl;
            const float *lrow = L.buf + (sl - cs) * kFp2Wcap;
            const float *grow = base + (size_t)sl * pitch;
#pragma unroll
            for (int j = 0; j < kFp2Rays; ++j) {
                if (!on[j] || sl < R[j].lo || sl > R[j].hi) continue;
                const float t = clamp_pad(fmaf(R[j].a, sf, R[j].c), hi_t);
                const float f = floorf(t);
                float2 q;
                if (staged) {
                    // inside the staged interval by its margin; the clamp
                    // only keeps the LDS address inside the row
                    const int o = min(max((int)f - t0, 0), kFp2Wcap - 2);
                    q = make_float2(lrow[o], lrow[o + 1]);
                } else {
                    q = ld2b(grow, (unsigned)f << 2);
                }
                const float x = lerp(q.x, q.y, t - f);
                if ((sl - R[j].lo) & 1)
                    acc1[j] += x;
                else
                    acc0[j] += x;
            }
        }
    }
#pragma unroll
    for (int j = 0; j < kFp2Rays; ++j)
        if (on[j]) val[j] = (acc0[j] + acc1[j]) * R[j].scale;
}

// ---- back projection with sinogram footprints staged in LDS -------------
// A workgroup of 256 threads covers 64 columns x (4 PY) rows (thread: column
// X0 + (tid & 63), rows Y0 + (tid >> 6) PY ..).  For every block of kBp2KA
// angles, the texels of each angle's sinogram row that the tile's taps can
// reach are copied into LDS with coalesced loads, and the taps read LDS.
// The reach is the bounding interval of the tap coordinate at the tile's
// four corner pixels (t is linear-fractional in the pixel position, so its
// extremes over the tile are at the corners, and the clamp is monotone),
// widened by one texel on each side, far beyond the fp32 rounding of the
// taps' incremental coordinates.  An angle whose interval exceeds
// kBp2Wcap texels, or whose denominator changes sign over the tile, is
// sampled from global memory.  Taps, weights and the order of the sums are
// those of bp2d_column, so both give the same floats.
constexpr int kBp2KA = 16;     // angles per staged block
constexpr int kBp2Wcap = 256;  // texels per staged angle row (power of two)

struct Bp2Lds {
    float buf[kBp2KA * kBp2Wcap];
    int2 fp[kBp2KA];  // (first texel, width); width 0: from global memory
};

template <int KIND>
__device__ __forceinline__ int2 bp2d_footprint(const BP2DAngle &c, float fx0, float fx1, float fy0,
                                               float fy1, float hi, int pitch) {
    float tmin = INFINITY, tmax = -INFINITY, dmin = INFINITY, dmax = -INFINITY;
    for (int j = 0; j < 4; ++j) {
        const float fx = (j & 1) ? fx1 : fx0, fy = (j & 2) ? fy1 : fy0;
        const float num = fmaf(fy, c.num[2], fmaf(fx, c.num[1], c.num[0]));
        const float den = fmaf(fy, c.den[2], fmaf(fx, c.den[1], c.den[0]));
        dmin = fminf(dmin, den);
        dmax = fmaxf(dmax, den);
        const float fr = (KIND == 0 || KIND == 3) ? 1.f : 1.f / den;
        const float t = clamp_pad(num * fr, hi);
        tmin = fminf(tmin, t);
        tmax = fmaxf(tmax, t);
    }
    const bool sign_ok = KIND == 0 || KIND == 3 || dmin > 0.f || dmax < 0.f;
    if (!sign_ok || !__builtin_isfinite(tmin) || !__builtin_isfinite(tmax)) return make_int2(0, 0);
    // taps read texels floor(t) and floor(t) + 1 (t >= 0.5 after the clamp)
    const int t0 = max(0, (int)tmin - 1), t1 = min(pitch - 1, (int)tmax + 2);
    const int wd = t1 - t0 + 1;
    return wd <= kBp2Wcap ? make_int2(t0, wd) : make_int2(0, 0);
}

template <int KIND, int PY>
__device__ __forceinline__ void bp2d_tile_lds(int X0, int Yt0, unsigned w, unsigned h, unsigned na,
                                              unsigned nd, const float *__restrict__ sp,
                                              const BP2DAngle *__restrict__ ang, float *acc,
                                              Bp2Lds &L) {
    const unsigned tid = threadIdx.x;
    const int X = X0 + (int)(tid & 63);
    const int Y0 = Yt0 + (int)(tid >> 6) * PY;
    const bool active = X < (int)w && Y0 < (int)h;
    const float fx = X - 0.5f * w + 0.5f;
    const float fy0 = -(Y0 - 0.5f * h + 0.5f);  // world y of row Y0
    // tile corners (pixels inside the slice)
    const float cx0 = X0 - 0.5f * w + 0.5f;
    const float cx1 = min(X0 + 63, (int)w - 1) - 0.5f * w + 0.5f;
    const float cy0 = -(Yt0 - 0.5f * h + 0.5f);
    const float cy1 = -(min(Yt0 + 4 * PY - 1, (int)h - 1) - 0.5f * h + 0.5f);
    const unsigned pitch = nd + 2 * kPad;
    const float hi = (float)nd + 2.5f;
    const __amdgpu_buffer_rsrc_t rs = texel_rsrc(sp);  // unstaged angles
#pragma unroll
    for (int i = 0; i < PY; ++i) acc[i] = 0.f;
    for (unsigned a0 = 0; a0 < na; a0 += kBp2KA) {
        __syncthreads();  // the previous block's rows are no longer read
        if (tid < (unsigned)kBp2KA)
            L.fp[tid] = (a0 + tid < na) ? bp2d_footprint<KIND>(ang[a0 + tid], cx0, cx1, cy0, cy1,
                                                               hi, (int)pitch)
                                        : make_int2(0, 0);
        __syncthreads();
#pragma unroll 4
        for (int i = (int)tid; i < kBp2KA * kBp2Wcap; i += 256) {
            const int k = i / kBp2Wcap, j = i & (kBp2Wcap - 1);
            const int2 f = L.fp[k];
            if (j < f.y) L.buf[i] = sp[(size_t)(a0 + k) * pitch + f.x + j];
        }
        __syncthreads();
        if (!active) continue;
        const unsigned nk = min((unsigned)kBp2KA, na - a0);
        for (unsigned k = 0; k < nk; ++k) {
            const BP2DAngle c = ang[a0 + k];
            // the footprint is the same for every lane: keep it in scalar
            // registers (a scalar branch, and the LDS row offset folded
            // into the tap's address computation)
            const int fx0 = __builtin_amdgcn_readfirstlane(L.fp[k].x);
            const int fwd = __builtin_amdgcn_readfirstlane(L.fp[k].y);
            float num = fmaf(fy0, c.num[2], fmaf(fx, c.num[1], c.num[0]));
            float den = fmaf(fy0, c.den[2], fmaf(fx, c.den[1], c.den[0]));
            const unsigned rowi = (a0 + k) * pitch;  // first texel of the row
            // LDS byte offset of texel 0 of the row: a tap's address is
            // one shift-and-add of its texel index
            const unsigned lb = __builtin_amdgcn_readfirstlane((unsigned)((int)k * kBp2Wcap - fx0) * 4u);
#pragma unroll
            for (int i = 0; i < PY; ++i) {
                float fr = (KIND == 0 || KIND == 3) ? 1.f : __builtin_amdgcn_rcpf(den);
                float t = clamp_pad(num * fr, hi);
                const int ti = (int)t;  // t >= 0.5: truncation is the floor
                float2 q;
                if (fwd) {
                    // texels ti, ti + 1 lie inside the staged interval
                    // [fx0, fx0 + fwd - 1]: the footprint widens the taps'
                    // range by a texel on each side, far beyond the rounding
                    // between its corner coordinates and the taps'
                    // incremental ones (the tests replay both on the host)
                    const float *l = reinterpret_cast<const float *>(
                        reinterpret_cast<const char *>(L.buf) + (lb + ((unsigned)ti << 2)));
                    q = make_float2(l[0], l[1]);
                } else {
                    q = ld2i(rs, rowi + (unsigned)ti);
                }
                float val = lerp(q.x, q.y, frac(t));
                float wgt = (KIND == 0) ? c.w : (KIND == 1 ? fr : (KIND == 2 ? fr * fr : 1.f));
                acc[i] = fmaf(wgt, val, acc[i]);
                num -= c.num[2];
                den -= c.den[2];
            }
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

// The PY outputs of a thread's pixel column: rows Y0 .. Y0 + nrows - 1
// (nrows <= PY) at vol[idx0 + i * stride], val = acc[i] * scale, stored as
// store_bp does.  Every load (old values, weights) is issued before the
// first store, so the column costs one memory round trip, not one per row
// (the compiler cannot move a load of vol above a store to vol by itself).
template <int PY>
__device__ __forceinline__ void store_bp_rows(float *vol, size_t idx0, size_t stride, int nrows,
                                              const float (&acc)[PY], float scale, int mode,
                                              const float *uw, float lo, float hi) {
    const bool old = mode != OUT_SET;
    const bool wts = (mode == OUT_EMUPD || mode == OUT_UPDATE) && uw;
    float ov[PY], wv[PY];
#pragma unroll
    for (int i = 0; i < PY; ++i) {
        ov[i] = 0.f;
        wv[i] = 1.f;
        if (i < nrows) {
            if (old) ov[i] = vol[idx0 + i * stride];
            if (wts) wv[i] = uw[idx0 + i * stride];
        }
    }
#pragma unroll
    for (int i = 0; i < PY; ++i) {
        if (i < nrows) {
            const float val = acc[i] * scale;
            float out;
            if (mode == OUT_ADD) out = ov[i] + val;
            else if (mode == OUT_SET) out = val;
            else if (mode == OUT_EMUPD) out = ov[i] * (val * wv[i]);
            else out = fminf(fmaxf(fmaf(wv[i], val, ov[i]), lo), hi);  // OUT_UPDATE
            vol[idx0 + i * stride] = out;
        }
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
