// Device helpers shared by the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace amx {

// Zero border (in texels) around every padded copy.  Kernels work in padded
// texel coordinates p = t + 2 (t = texture coordinate - 0.5) clamped to
// [0.5, n + 2.5]: the two taps of a linear fetch then land in [0, n + 3],
// inside the array plus its 2-texel zero border.  The clamp never changes a
// value (both taps outside the data read 0, as with border addressing) and
// every access stays in bounds.  v_med3_f32 returns a finite value or NaN;
// a NaN coordinate converts to offset 0, which is also in bounds.
constexpr int kPad = 2;

__device__ __forceinline__ float clamp_pad(float p, float hi) {
    return __builtin_amdgcn_fmed3f(p, 0.5f, hi);
}

// Pair of floats at base + byte offset (one global_load_dwordx2 with a
// scalar base and a 32-bit vector offset).
__device__ __forceinline__ float2 ld2b(const float *base, unsigned byte_off) {
    typedef float f2u __attribute__((ext_vector_type(2), aligned(4)));
    f2u v = *reinterpret_cast<const f2u *>(reinterpret_cast<const char *>(base) + byte_off);
    return make_float2(v.x, v.y);
}

// Four floats at base + byte offset (one global_load_dwordx4; 8-byte
// aligned): the bilinear taps of a row-pair interleaved plane.
__device__ __forceinline__ float4 ld4b(const float *base, unsigned byte_off) {
    typedef float f4u __attribute__((ext_vector_type(4), aligned(8)));
    f4u v = *reinterpret_cast<const f4u *>(reinterpret_cast<const char *>(base) + byte_off);
    return make_float4(v.x, v.y, v.z, v.w);
}

// Wave-uniform copy of a pointer (from the first lane): keeps pointer
// arithmetic on it in scalar registers.
template <class T>
__device__ __forceinline__ T *first_lane_ptr(T *p) {
    const unsigned long long u = (unsigned long long)p;
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)u);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(u >> 32));
    return (T *)(((unsigned long long)hi << 32) | lo);
}

// Buffer resources (V#) for loads whose base is wave-uniform: the base
// goes into scalar registers and the texture unit adds the per-lane
// record index x stride (and a scalar byte offset), so a load costs one
// vector register of address and no 64-bit vector address arithmetic.
// num_records = 0xffffffff (no range check: the callers' indices stay in
// their arrays); word 3 = 32-bit data format, as gfx9 buffer loads expect.
// The loads are the structured-buffer intrinsic (index, offset, soffset,
// aux), declared by its LLVM name.
//
// Structured buffer of 8-byte records (row-difference pairs): a 16-byte
// load of records index and index + 1 at byte offset soff; the index is
// one v_mad_u32_u24 from the texel coordinates.
typedef unsigned amx_u4 __attribute__((ext_vector_type(4)));
__device__ amx_u4 amx_struct_load_v4(__amdgpu_buffer_rsrc_t r, int vindex, int voffset,
                                     int soffset, int aux)
    __asm("llvm.amdgcn.struct.ptr.buffer.load.v4i32");

__device__ __forceinline__ __amdgpu_buffer_rsrc_t pair_rsrc(const void *base) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(first_lane_ptr(base)), (short)8,
                                             -1, 0x00020000);
}

// Structured buffer of 4-byte records (a padded 2D slice copy) from a
// wave-uniform base: 8- and 16-byte loads of consecutive texels from the
// record index, the only per-lane address term.
typedef unsigned amx_u2 __attribute__((ext_vector_type(2)));
__device__ amx_u2 amx_struct_load_v2(__amdgpu_buffer_rsrc_t r, int vindex, int voffset,
                                     int soffset, int aux)
    __asm("llvm.amdgcn.struct.ptr.buffer.load.v2i32");

__device__ __forceinline__ __amdgpu_buffer_rsrc_t texel_rsrc(const float *base) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(first_lane_ptr(base)), (short)4,
                                             -1, 0x00020000);
}

// texels index and index + 1
__device__ __forceinline__ float2 ld2i(__amdgpu_buffer_rsrc_t r, unsigned index) {
    typedef float f2 __attribute__((ext_vector_type(2)));
    const f2 v = __builtin_bit_cast(f2, amx_struct_load_v2(r, (int)index, 0, 0, 0));
    return make_float2(v.x, v.y);
}

// texels index .. index + 3
__device__ __forceinline__ float4 ld4i(__amdgpu_buffer_rsrc_t r, unsigned index) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    const f4 v = __builtin_bit_cast(f4, amx_struct_load_v4(r, (int)index, 0, 0, 0));
    return make_float4(v.x, v.y, v.z, v.w);
}

__device__ __forceinline__ float4 ld4p(__amdgpu_buffer_rsrc_t r, unsigned index, unsigned soff) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    const f4 v = __builtin_bit_cast(f4, amx_struct_load_v4(r, (int)index, 0, (int)soff, 0));
    return make_float4(v.x, v.y, v.z, v.w);
}

// Four floats at base + byte offset, 4-byte aligned (one
// global_load_dwordx4): two linear taps of one row when their texels lie
// within four.
__device__ __forceinline__ float4 ld4u(const float *base, unsigned byte_off) {
    typedef float f4u __attribute__((ext_vector_type(4), aligned(4)));
    f4u v = *reinterpret_cast<const f4u *>(reinterpret_cast<const char *>(base) + byte_off);
    return make_float4(v.x, v.y, v.z, v.w);
}

// Texels (g + d, g + d + 1) out of four loaded from texel g, d in {0, 1, 2}
// (a float holding an integer).
__device__ __forceinline__ float2 pick2(float4 x, float d) {
    const bool ge1 = d >= 0.5f, ge2 = d >= 1.5f;
    return make_float2(ge1 ? (ge2 ? x.z : x.y) : x.x, ge1 ? (ge2 ? x.w : x.z) : x.y);
}

// 16-byte aligned float4 / one float at base + byte offset (scalar base,
// 32-bit vector offset)
__device__ __forceinline__ float4 ld4a(const float4 *base, unsigned byte_off) {
    return *reinterpret_cast<const float4 *>(reinterpret_cast<const char *>(base) + byte_off);
}
__device__ __forceinline__ float ld1b(const float *base, unsigned byte_off) {
    return *reinterpret_cast<const float *>(reinterpret_cast<const char *>(base) + byte_off);
}

// Byte offset of the tap pair of padded coordinates (p1, p2) in a
// row-pair plane of `pitch` pairs per row (8 bytes per pair).  p >= 0.5
// after clamp_pad, so truncation (v_cvt_u32_f32) is the floor; NaN
// converts to 0, which is in bounds.
__device__ __forceinline__ unsigned pair_off(float p1, float p2, unsigned pitch) {
    return (__umul24((unsigned)p2, pitch) + (unsigned)p1) << 3;
}

// the same as a record index (8-byte records)
__device__ __forceinline__ unsigned pair_idx(float p1, float p2, unsigned pitch) {
    return __umul24((unsigned)p2, pitch) + (unsigned)p1;
}

// fractional part of a clamped (>= 0.5) padded coordinate: p - floor(p)
__device__ __forceinline__ float frac(float p) { return __builtin_amdgcn_fractf(p); }

__device__ __forceinline__ float lerp(float a, float b, float w) { return fmaf(w, b - a, a); }

// Bilinear value from the load of two difference pairs: q = (v[r][c],
// v[r+1][c] - v[r][c], v[r][c+1], v[r+1][c+1] - v[r][c+1]); w1 along c, w2
// along r.  The stored differences are the ones lerp() would compute, so
// each row-direction lerp is one fma.
__device__ __forceinline__ float bilerp_pairs(float4 q, float w1, float w2) {
    const float l0 = fmaf(w2, q.y, q.x), l1 = fmaf(w2, q.w, q.z);
    return fmaf(w1, l1 - l0, l0);
}

__device__ __forceinline__ int wave_min(int v) {
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ int wave_max(int v) {
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}

// Tent weight (1 - |d|)+ of the linear / bilinear interpolation.  Written
// as a clamp to [0, 1] (1 - |d| <= 1, so the upper bound changes nothing and
// NaN still gives 0), which folds into the subtraction's clamp bit: one
// instruction instead of a subtraction and a maximum.
__device__ __forceinline__ float tent(float d) {
    return fminf(fmaxf(1.f - fabsf(d), 0.f), 1.f);
}

// Slice range [lo, hi] of s in [0, n0) for which -1 < a*s + c < n can hold
// (widened by one slice on each side; exactness is not required because
// samples outside the array read the zero border).
__device__ __forceinline__ void clip_range(double a, double c, double n, int n0, int &lo, int &hi) {
    if (fabs(a) < 1e-12) {
        if (!(c > -1.0 && c < n)) { lo = 1; hi = 0; }
        return;
    }
    double s1 = (-1.0 - c) / a, s2 = (n - c) / a;
    double smin = fmin(s1, s2), smax = fmax(s1, s2);
    double l = floor(smin) - 1.0, h = ceil(smax) + 1.0;
    if (l > (double)lo) lo = (l > (double)n0) ? n0 : (int)l;
    if (h < (double)hi) hi = (h < -1.0) ? -1 : (int)h;
}

// Output epilogue modes (internal)
// OUT_ADD    out += v
// OUT_SET    out  = v
// OUT_RESID  out  = (b + v) * w            (w optional)
// OUT_UPDATE out  = clamp(out + w * v)     (w optional)
// OUT_ADDMUL out  = (out + v) * w          (w optional)
// OUT_RATIO  out  = v > 1e-6 ? b / v : 0   (FP; EM's sinogram ratio)
// OUT_EMUPD  out  = out * (v * w)          (BP; EM's multiplicative update)
enum { OUT_ADD = 0, OUT_SET = 1, OUT_RESID = 2, OUT_UPDATE = 3, OUT_ADDMUL = 4, OUT_RATIO = 5,
       OUT_EMUPD = 6 };

// EM's sinogram ratio (opDividedBy, cuda/2d/arith.cu:91-98)
__device__ __forceinline__ float em_ratio(float b, float v) { return v > 0.000001f ? b / v : 0.f; }

}  // namespace amx
