// 3D projectors for gfx950: cone / par3d ray-driven forward projection and
// voxel-driven back projection (default, FDK and 2D weighting), plus the
// exact-transpose (matched) back projection.
//
// Data layout in HBM
//   user volume       [z][y][x]
//   user projections  [v][angle][u]
//   FP input copies   slice-major and zero-padded per direction group:
//                       x-group  [x][z+4][y+4]   (c0 = x, c2 = z, c1 = y)
//                       y-group  [y][z+4][x+4]   (c0 = y, c2 = z, c1 = x)
//                       z-group  [z][y+4][x+4]   (c0 = z, c2 = y, c1 = x)
//                     so that the in-slice fast axis is contiguous for the
//                     rays of every group and one slice is one contiguous
//                     plane (32-bit offsets inside a plane).
//   BP input copy     angle-major and zero-padded: [angle][v+4][u+4]
//   Both copies are stored row-pair interleaved: element [r][c] is the
//   float2 (p[r][c], p[r+1][c]), so the four taps of a bilinear sample at
//   (r, c) are one 16-byte load of [r][c] and [r][c+1].  The texture
//   address unit handles one instruction per sample instead of two (it is
//   the unit that limits both kernels); the copies take twice the memory.
//
// Sampling semantics follow cuda/3d/cone_fp.cu:154-231,
// cuda/3d/par3d_fp.cu:152-230, cuda/3d/cone_bp.cu:55-133 and
// cuda/3d/par3d_bp.cu:55-120 with ideal fp32 (bi)linear interpolation in
// place of the 8-bit-weight texture unit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>

#include "devutil.hpp"
#include "internal.hpp"

namespace amx {

// ----------------------------------------------------------- prep --------
// dst[c0][c2 + 2][c1 + 2] = (vol at c2, vol at c2 + 1) for the z- and
// y-groups (the fast axis stays x; row pairs as described above).  One
// workgroup per destination row; borders written 0.
__global__ __launch_bounds__(256) void pad_copy_xfast(const float *__restrict__ vol,
                                                      float2 *__restrict__ dst, unsigned nx,
                                                      unsigned ny, unsigned nz, int group) {
    const unsigned px = nx + 2 * kPad;
    const unsigned row = blockIdx.x;  // c0 * n2p + padded c2
    const unsigned n2p = (group == 2) ? ny + 2 * kPad : nz + 2 * kPad;
    const int n2 = (group == 2) ? (int)ny : (int)nz;
    const unsigned c0 = row / n2p;
    const int c2 = (int)(row % n2p) - kPad;
    float2 *out = dst + (size_t)row * px;
    const bool in0 = c2 >= 0 && c2 < n2, in1 = c2 + 1 >= 0 && c2 + 1 < n2;
    auto src_row = [&](int r) {
        return (group == 2) ? vol + ((size_t)c0 * ny + r) * nx     // [z=c0][y=r]
                            : vol + ((size_t)r * ny + c0) * nx;    // [z=r][y=c0]
    };
    const float *s0 = src_row(in0 ? c2 : 0), *s1 = src_row(in1 ? c2 + 1 : 0);
    for (unsigned i = threadIdx.x; i < px; i += blockDim.x) {
        const int x = (int)i - kPad;
        const bool xin = x >= 0 && x < (int)nx;
        out[i] = make_float2((in0 && xin) ? s0[x] : 0.f, (in1 && xin) ? s1[x] : 0.f);
    }
}

// x-group: dst[x][z + 2][y + 2] = (vol[z][y][x], vol[z+1][y][x]); per
// padded z row an (x, y) transpose of planes z and z + 1 through two
// 64 x 64 LDS tiles.  grid = (y tiles, x tiles, nz + 4).
__global__ __launch_bounds__(256) void pad_transpose_x(const float *__restrict__ vol,
                                                       float2 *__restrict__ dst, unsigned nx,
                                                       unsigned ny, unsigned nz) {
    __shared__ float tile[2][64][65];
    const unsigned pyn = ny + 2 * kPad, pzn = nz + 2 * kPad;
    const int zr = blockIdx.z;  // padded z row
    const int x0 = blockIdx.y * 64;
    const int yp0 = blockIdx.x * 64;  // padded y tile origin
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int k = 0; k < 2; ++k) {
        const int z = zr - kPad + k;
        for (int r = ty; r < 64; r += 4) {
            int y = yp0 + r - kPad;
            int x = x0 + tx;
            float v = 0.f;
            if (z >= 0 && z < (int)nz && y >= 0 && y < (int)ny && x < (int)nx)
                v = vol[((size_t)z * ny + y) * nx + x];
            tile[k][r][tx] = v;
        }
    }
    __syncthreads();
    for (int r = ty; r < 64; r += 4) {
        int x = x0 + r;
        int yp = yp0 + tx;
        if (x < (int)nx && yp < (int)pyn)
            dst[((size_t)x * pzn + zr) * pyn + yp] = make_float2(tile[0][tx][r], tile[1][tx][r]);
    }
}

// proj [v][a][u] -> padded angle-major [a][v + 2][u + 2] (borders 0), the
// single (not paired) layout of the algorithms' projection buffers.
__global__ __launch_bounds__(256) void proj_to_am(const float *__restrict__ proj,
                                                  float *__restrict__ pp, unsigned nu,
                                                  unsigned na, unsigned nv) {
    const unsigned pu = nu + 2 * kPad, pv = nv + 2 * kPad;
    const unsigned row = blockIdx.x;  // a * pv + padded v
    const unsigned a = row / pv;
    const int v = (int)(row % pv) - kPad;
    float *out = pp + (size_t)row * pu;
    const bool inside = v >= 0 && v < (int)nv;
    const float *src = proj + ((size_t)(inside ? v : 0) * na + a) * nu;
    for (unsigned i = threadIdx.x; i < pu; i += blockDim.x) {
        int u = (int)i - kPad;
        out[i] = (inside && u >= 0 && u < (int)nu) ? src[u] : 0.f;
    }
}

// proj [v][a][u] -> padded angle-major row pairs [a][v + 2][u + 2] =
// (p[v][u], p[v+1][u]) (borders 0).  One workgroup per destination row.
__global__ __launch_bounds__(256) void proj_to_am_pairs(const float *__restrict__ proj,
                                                        float2 *__restrict__ pp, unsigned nu,
                                                        unsigned na, unsigned nv) {
    const unsigned pu = nu + 2 * kPad, pv = nv + 2 * kPad;
    const unsigned row = blockIdx.x;  // a * pv + padded v
    const unsigned a = row / pv;
    const int v = (int)(row % pv) - kPad;
    float2 *out = pp + (size_t)row * pu;
    const bool in0 = v >= 0 && v < (int)nv, in1 = v + 1 >= 0 && v + 1 < (int)nv;
    const float *s0 = proj + ((size_t)(in0 ? v : 0) * na + a) * nu;
    const float *s1 = proj + ((size_t)(in1 ? v + 1 : 0) * na + a) * nu;
    for (unsigned i = threadIdx.x; i < pu; i += blockDim.x) {
        const int u = (int)i - kPad;
        const bool uin = u >= 0 && u < (int)nu;
        out[i] = make_float2((in0 && uin) ? s0[u] : 0.f, (in1 && uin) ? s1[u] : 0.f);
    }
}

// padded single layout [a][v + 4][u + 4] -> row pairs (same padding).
__global__ __launch_bounds__(256) void am_to_pairs(const float *__restrict__ src,
                                                   float2 *__restrict__ pp, unsigned nu,
                                                   unsigned nv) {
    const unsigned pu = nu + 2 * kPad, pv = nv + 2 * kPad;
    const unsigned row = blockIdx.x;  // a * pv + padded v
    const bool last = (row % pv) == pv - 1;
    const float *s0 = src + (size_t)row * pu;
    float2 *out = pp + (size_t)row * pu;
    for (unsigned i = threadIdx.x; i < pu; i += blockDim.x)
        out[i] = make_float2(s0[i], last ? 0.f : s0[pu + i]);
}

// ------------------------------------------------------------- FP --------
struct FPArgs {
    const float *volp;  // padded slice-major copy of this group
    unsigned n0, n1, n2;
    const FP3DAngle *ang;
    const int *alist;  // angle indices of this direction group
    unsigned nu, nv, na;
    float fs1, fs2, osc;
    float *out;
    int mode;
    const float *rb;  // OUT_RESID: b in the output layout
    const float *rw;  // OUT_RESID: weights in the output layout (may be NULL)
    int padded_out;   // write [a][v+2][u+2] instead of [v][a][u]
    // workgroup map (fp_map): detector tiles, angles per block, and the
    // range [lfirst, lfirst + lcount) of alist covered by this launch
    unsigned tiles, ablk, lfirst, lcount;
    unsigned ss;  // rays per detector pixel and dimension (DetectorSuperSampling)
};

// Workgroup -> (detector tile, angle) map.  Workgroups are handed out
// round-robin over the 8 XCDs (workgroup L runs on XCD L % 8), and each XCD
// has its own L2.  The grid is a multiple of 8; XCD x takes the contiguous
// logical range [x * per, (x + 1) * per), in the order (angle block, tile,
// angle within block).  The workgroups resident on one XCD at a time are
// then a few neighbouring tiles, each over `ablk` consecutive angles, which
// start together and march through the same slices at the same time: rays
// of neighbouring angles sample nearly the same voxels, so they share the
// XCD's L2 instead of each fetching the volume from HBM.
__device__ __forceinline__ bool fp_map(const FPArgs &A, unsigned &tile, int &angle) {
    const unsigned L = blockIdx.x;
    const unsigned per = gridDim.x >> 3;
    const unsigned logical = (L & 7u) * per + (L >> 3);
    const unsigned a_in = logical % A.ablk;
    const unsigned rest = logical / A.ablk;
    tile = rest % A.tiles;
    const unsigned li = (rest / A.tiles) * A.ablk + a_in;
    if (li >= A.lcount) return false;
    angle = A.alist[A.lfirst + li];
    return true;
}

// One thread per ray; a 256-thread workgroup covers a 16 x 16 detector tile
// of one angle, each wave a 16 u x 4 v sub-tile.  Ray setup runs in FP64;
// the slice loop is wave-uniform over the union of the lanes' clipped slice
// ranges, so the slice base address is scalar and every fetch is one
// dwordx2 load with a 32-bit vector offset.
// (u, v): detector coordinates of the ray relative to pixel (0, 0)'s centre
// (fractional for sub-rays).
template <bool CONE>
__device__ __forceinline__ void fp3d_ray(const FPArgs &A, const FP3DAngle &G, double u, double v,
                                         float &a1f, float &a2f, float &c1f, float &c2f,
                                         float &scale, int &lo, int &hi) {
    double P[3], d[3];
    for (int k = 0; k < 3; ++k) P[k] = G.p0[k] + u * G.u[k] + v * G.v[k];
    for (int k = 0; k < 3; ++k) d[k] = CONE ? (G.s[k] - P[k]) : G.s[k];
    const double *B = CONE ? G.s : P;
    double a1 = d[1] / d[0], a2 = d[2] / d[0];
    // texel coordinate - 0.5 at slice s: a_k (s - n0/2 + 0.5 - B0) + B_k + n_k/2 - 0.5
    double base = -0.5 * A.n0 + 0.5 - B[0];
    double c1 = a1 * base + B[1] + 0.5 * A.n1 - 0.5;
    double c2 = a2 * base + B[2] + 0.5 * A.n2 - 0.5;
    lo = 0;
    hi = (int)A.n0 - 1;
    clip_range(a1, c1, (double)A.n1, (int)A.n0, lo, hi);
    clip_range(a2, c2, (double)A.n2, (int)A.n0, lo, hi);
    a1f = (float)a1;
    a2f = (float)a2;
    c1f = (float)(c1 + kPad);  // padded texel coordinates
    c2f = (float)(c2 + kPad);
    scale = (float)(sqrt(a1 * a1 * A.fs1 + a2 * a2 * A.fs2 + 1.0) * A.osc);
}

// Sum of the bilinear samples of slices [slo, shi] (wave-uniform) along the
// ray (c1, c2) = a s + c through the padded group copy.
__device__ __forceinline__ float march3d(const FPArgs &A, float a1f, float a2f, float c1f,
                                         float c2f, int slo, int shi) {
    const unsigned pitch = A.n1 + 2 * kPad;                        // pairs per row
    const size_t plane = (size_t)(A.n2 + 2 * kPad) * pitch * 2;     // floats per slice
    const float pitchf = (float)pitch;
    const float hi1 = (float)A.n1 + 2.5f, hi2 = (float)A.n2 + 2.5f;
    float acc0 = 0.f, acc1 = 0.f;
    int s = slo;
    const float *pl = A.volp + (size_t)(slo > 0 ? slo : 0) * plane;
    for (; s + 1 <= shi; s += 2) {
        const float sf = (float)s;
        float p1 = clamp_pad(fmaf(a1f, sf, c1f), hi1);
        float p2 = clamp_pad(fmaf(a2f, sf, c2f), hi2);
        float q1 = clamp_pad(fmaf(a1f, sf + 1.f, c1f), hi1);
        float q2 = clamp_pad(fmaf(a2f, sf + 1.f, c2f), hi2);
        float fp1 = floorf(p1), fp2 = floorf(p2), fq1 = floorf(q1), fq2 = floorf(q2);
        float4 r = ld4b(pl, pair_off(fp1, fp2, pitchf));
        float4 t = ld4b(pl + plane, pair_off(fq1, fq2, pitchf));
        acc0 += bilerp_pairs(r, p1 - fp1, p2 - fp2);
        acc1 += bilerp_pairs(t, q1 - fq1, q2 - fq2);
        pl += 2 * plane;
    }
    if (s <= shi) {
        const float sf = (float)s;
        float p1 = clamp_pad(fmaf(a1f, sf, c1f), hi1);
        float p2 = clamp_pad(fmaf(a2f, sf, c2f), hi2);
        float fp1 = floorf(p1), fp2 = floorf(p2);
        acc0 += bilerp_pairs(ld4b(pl, pair_off(fp1, fp2, pitchf)), p1 - fp1, p2 - fp2);
    }
    return acc0 + acc1;
}

// SS: A.ss x A.ss sub-rays per detector pixel at offsets (k + 0.5) / ss - 0.5
// in u and v, each with its own distance weight, averaged
// (cone_FP_SS_t / par3D_FP_SS_t, cuda/3d/cone_fp.cu:233-323,
// cuda/3d/par3d_fp.cu:234-325).
template <bool CONE, bool SS>
__global__ __launch_bounds__(256) void fp3d_kernel(FPArgs A) {
    unsigned tile;
    int angle;
    if (!fp_map(A, tile, angle)) return;  // uniform over the workgroup
    const unsigned tiles_u = (A.nu + 15) / 16;
    const unsigned tu = tile % tiles_u, tv = tile / tiles_u;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int u = tu * 16 + (lane & 15);
    const int v = tv * 16 + wave * 4 + (lane >> 4);
    const FP3DAngle &G = A.ang[angle];
    const bool active = u < (int)A.nu && v < (int)A.nv;

    float val = 0.f;
    if (!SS) {
        float a1f = 0.f, a2f = 0.f, c1f = 0.5f, c2f = 0.5f, scale = 0.f;
        int lo = (int)A.n0, hi = -1;
        if (active) fp3d_ray<CONE>(A, G, u, v, a1f, a2f, c1f, c2f, scale, lo, hi);
        const int slo = __builtin_amdgcn_readfirstlane(wave_min(lo));
        const int shi = __builtin_amdgcn_readfirstlane(wave_max(hi));
        val = march3d(A, a1f, a2f, c1f, c2f, slo, shi) * scale;
    } else {
        const int R = (int)A.ss;
        for (int k = 0; k < R; ++k)
            for (int l = 0; l < R; ++l) {
                float a1f = 0.f, a2f = 0.f, c1f = 0.5f, c2f = 0.5f, scale = 0.f;
                int lo = (int)A.n0, hi = -1;
                if (active)
                    fp3d_ray<CONE>(A, G, u - 0.5 + (k + 0.5) / R, v - 0.5 + (l + 0.5) / R, a1f,
                                   a2f, c1f, c2f, scale, lo, hi);
                const int slo = __builtin_amdgcn_readfirstlane(wave_min(lo));
                const int shi = __builtin_amdgcn_readfirstlane(wave_max(hi));
                val += march3d(A, a1f, a2f, c1f, c2f, slo, shi) * scale;
            }
        val /= (float)(R * R);
    }
    if (!active) return;
    size_t idx = ((size_t)v * A.na + angle) * A.nu + u;
    size_t oidx = A.padded_out
                      ? ((size_t)angle * (A.nv + 2 * kPad) + (v + kPad)) * (A.nu + 2 * kPad) + u + kPad
                      : idx;
    if (A.mode == OUT_ADD) {
        A.out[oidx] += val;
    } else if (A.mode == OUT_SET) {
        A.out[oidx] = val;
    } else if (A.mode == OUT_RESID) {  // (b + val) * w
        float r = A.rb[idx] + val;
        if (A.rw) r *= A.rw[oidx];
        A.out[oidx] = r;
    } else if (A.mode == OUT_RATIO) {
        A.out[oidx] = em_ratio(A.rb[idx], val);
    } else {  // OUT_ADDMUL: (out + val) * w
        float r = A.out[oidx] + val;
        if (A.rw) r *= A.rw[oidx];
        A.out[oidx] = r;
    }
}

// Squared weights of the par3d FP: per ray, the sum over slices of the
// squared bilinear weights of its samples (separable: per axis
// sum w^2 = t^2 + (1-t)^2, or the one tap inside near the border), times
// the squared distance weight (par3D_FP_SumSqW_t and dirWeights,
// cuda/3d/par3d_fp.cu:329-419).  No volume data is read.
__device__ __forceinline__ float dir_weight_sq(float f, float n) {
    // f: texture coordinate (texel i covers [i, i + 1))
    if (f <= -0.5f) return 0.f;
    if (f <= 0.5f) return (f + 0.5f) * (f + 0.5f);
    if (f <= n - 0.5f) {
        float t = f + 0.5f - floorf(f + 0.5f);
        return t * t + (1.f - t) * (1.f - t);
    }
    if (f <= n + 0.5f) return (n + 0.5f - f) * (n + 0.5f - f);
    return 0.f;
}

__global__ __launch_bounds__(256) void fp3d_sumsq_kernel(FPArgs A) {
    unsigned tile;
    int angle;
    if (!fp_map(A, tile, angle)) return;
    const unsigned tiles_u = (A.nu + 15) / 16;
    const unsigned tu = tile % tiles_u, tv = tile / tiles_u;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int u = tu * 16 + (lane & 15);
    const int v = tv * 16 + wave * 4 + (lane >> 4);
    if (u >= (int)A.nu || v >= (int)A.nv) return;
    float a1f, a2f, c1f, c2f, scale;
    int lo, hi;
    fp3d_ray<false>(A, A.ang[angle], u, v, a1f, a2f, c1f, c2f, scale, lo, hi);
    // padded texel coordinate p = f - 0.5 + kPad
    const float o1 = c1f - kPad + 0.5f, o2 = c2f - kPad + 0.5f;
    const float n1 = (float)A.n1, n2 = (float)A.n2;
    float acc = 0.f;
    for (int s = lo; s <= hi; ++s)
        acc += dir_weight_sq(fmaf(a1f, (float)s, o1), n1) * dir_weight_sq(fmaf(a2f, (float)s, o2), n2);
    const float val = acc * scale * scale;
    const size_t idx = ((size_t)v * A.na + angle) * A.nu + u;
    if (A.mode == OUT_ADD)
        A.out[idx] += val;
    else
        A.out[idx] = val;
}

// ------------------------------------------------------------- BP --------
struct BPArgs {
    const float *pp;  // padded angle-major projections
    unsigned nu, nv, na;
    const BP3DAngle *ang;
    const float *aw;  // par3d per-angle weights
    unsigned nx, ny, nz;
    float osc;
    float *vol;
    int mode;
    const float *uw;  // OUT_UPDATE: per-voxel weights (may be NULL)
    float lo, hi;     // OUT_UPDATE: clamp
};

// KIND: 0 cone fr^2 (default / FDK), 1 cone fr (2D weighting), 2 par3d.
// One thread per (x, y) column of ZS voxels; 16 x 16 threads per group.
// The angle loop is uniform, so the per-angle coefficients are scalar loads
// and the angle's projection plane base is a scalar address.
// ZC: U and the denominator are constant along z for every angle (checked
// on the host: nu[2] == dn[2] == 0); the reciprocal, the U coordinate, its
// tap offset and weight and the voxel weight are then computed once per
// column and angle.  Both forms take the same fp32 operations on the same
// values (adding a zero increment is exact), so they give identical results.
template <int KIND, int ZS, bool ZC>
__global__ __launch_bounds__(256) void bp3d_kernel(BPArgs A) {
    const unsigned tiles_x = (A.nx + 15) / 16;
    const int X = (blockIdx.x % tiles_x) * 16 + (threadIdx.x & 15);
    const int Y = (blockIdx.x / tiles_x) * 16 + (threadIdx.x >> 4);
    const int Z0 = blockIdx.y * ZS;
    if (X >= (int)A.nx || Y >= (int)A.ny) return;
    const float fx = X - 0.5f * A.nx + 0.5f;
    const float fy = Y - 0.5f * A.ny + 0.5f;
    const float fz = Z0 - 0.5f * A.nz + 0.5f;
    const unsigned pitch = A.nu + 2 * kPad;                    // pairs per row
    const size_t plane = (size_t)(A.nv + 2 * kPad) * pitch * 2; // floats per angle
    const float pitchf = (float)pitch;
    const unsigned rowb = pitch * 8u;                           // bytes per row
    const float hiu = (float)A.nu + 2.5f, hiv = (float)A.nv + 2.5f;
    float acc[ZS];
#pragma unroll
    for (int i = 0; i < ZS; ++i) acc[i] = 0.f;

    const float *pl = A.pp;
    for (unsigned a = 0; a < A.na; ++a, pl += plane) {
        const BP3DAngle c = A.ang[a];
        float un = fmaf(fz, c.nu[2], fmaf(fy, c.nu[1], fmaf(fx, c.nu[0], c.nu[3])));
        float vn = fmaf(fz, c.nv[2], fmaf(fy, c.nv[1], fmaf(fx, c.nv[0], c.nv[3])));
        float dn = fmaf(fz, c.dn[2], fmaf(fy, c.dn[1], fmaf(fx, c.dn[0], c.dn[3])));
        const float wa = (KIND == 2) ? A.aw[a] : 1.f;
        if (ZC) {
            const float fr = (KIND == 2) ? 1.f : __builtin_amdgcn_rcpf(dn);
            const float pu = clamp_pad(un * fr, hiu);
            const float fu = floorf(pu);
            const float wu = pu - fu;
            const unsigned ou = (unsigned)fu << 3;
            const float wgt = (KIND == 0) ? fr * fr : (KIND == 1 ? fr : wa);
#pragma unroll
            for (int i = 0; i < ZS; ++i) {
                float pv = clamp_pad(vn * fr, hiv);
                // pv in [0.5, nv + 2.5]: floor is exact and fract = pv - floor(pv)
                unsigned iv = (unsigned)floorf(pv);
                float wv = __builtin_amdgcn_fractf(pv);
                float val = bilerp_pairs(ld4b(pl, __umul24(iv, rowb) + ou), wu, wv);
                acc[i] = fmaf(wgt, val, acc[i]);
                vn += c.nv[2];
            }
        } else {
#pragma unroll
            for (int i = 0; i < ZS; ++i) {
                float fr = (KIND == 2) ? 1.f : __builtin_amdgcn_rcpf(dn);
                float pu = clamp_pad(un * fr, hiu);
                float pv = clamp_pad(vn * fr, hiv);
                float fu = floorf(pu), fv = floorf(pv);
                float val = bilerp_pairs(ld4b(pl, pair_off(fu, fv, pitchf)), pu - fu, pv - fv);
                float wgt = (KIND == 0) ? fr * fr : (KIND == 1 ? fr : wa);
                acc[i] = fmaf(wgt, val, acc[i]);
                un += c.nu[2];
                vn += c.nv[2];
                dn += c.dn[2];
            }
        }
    }
    const int nzr = min(ZS, (int)A.nz - Z0);
    for (int i = 0; i < ZS; ++i) {
        if (i >= nzr) break;
        size_t idx = ((size_t)(Z0 + i) * A.ny + Y) * A.nx + X;
        float val = acc[i] * A.osc;
        if (A.mode == OUT_ADD) {
            A.vol[idx] += val;
        } else if (A.mode == OUT_SET) {
            A.vol[idx] = val;
        } else if (A.mode == OUT_EMUPD) {
            A.vol[idx] *= val * A.uw[idx];
        } else {  // OUT_UPDATE: clamp(x + w * val)
            float w = A.uw ? A.uw[idx] : 1.f;
            A.vol[idx] = fminf(fmaxf(fmaf(w, val, A.vol[idx]), A.lo), A.hi);
        }
    }
}

// Supersampled BP (dev_cone_BP_SS / dev_par3D_BP_SS,
// cuda/3d/cone_bp.cu:138-217, cuda/3d/par3d_bp.cu:123-198): R^3 sub-voxels
// per voxel at offsets (k + 0.5) / R - 0.5, averaged.  KIND 0: cone, weight
// fr^2; KIND 2: par3d, the angle's weight.  One thread per voxel.
template <int KIND>
__global__ __launch_bounds__(256) void bp3d_ss_kernel(BPArgs A, int R) {
    const unsigned tiles_x = (A.nx + 15) / 16;
    const int X = (blockIdx.x % tiles_x) * 16 + (threadIdx.x & 15);
    const int Y = (blockIdx.x / tiles_x) * 16 + (threadIdx.x >> 4);
    const int Z = blockIdx.y;
    if (X >= (int)A.nx || Y >= (int)A.ny) return;
    const float sub = 1.0f / (float)R;
    const float fx0 = X - 0.5f * A.nx + 0.5f * sub;
    const float fy0 = Y - 0.5f * A.ny + 0.5f * sub;
    const float fz0 = Z - 0.5f * A.nz + 0.5f * sub;
    const unsigned pitch = A.nu + 2 * kPad;
    const size_t plane = (size_t)(A.nv + 2 * kPad) * pitch * 2;
    const float pitchf = (float)pitch;
    const float hiu = (float)A.nu + 2.5f, hiv = (float)A.nv + 2.5f;
    float acc = 0.f;
    const float *pl = A.pp;
    for (unsigned a = 0; a < A.na; ++a, pl += plane) {
        const BP3DAngle c = A.ang[a];
        const float wa = (KIND == 2) ? A.aw[a] : 1.f;
        for (int i = 0; i < R; ++i) {
            const float fx = fx0 + i * sub;
            for (int j = 0; j < R; ++j) {
                const float fy = fy0 + j * sub;
                for (int k = 0; k < R; ++k) {
                    const float fz = fz0 + k * sub;
                    float un = fmaf(fz, c.nu[2], fmaf(fy, c.nu[1], fmaf(fx, c.nu[0], c.nu[3])));
                    float vn = fmaf(fz, c.nv[2], fmaf(fy, c.nv[1], fmaf(fx, c.nv[0], c.nv[3])));
                    float fr = 1.f;
                    if (KIND != 2) {
                        float dn = fmaf(fz, c.dn[2], fmaf(fy, c.dn[1], fmaf(fx, c.dn[0], c.dn[3])));
                        fr = __builtin_amdgcn_rcpf(dn);
                    }
                    float pu = clamp_pad(un * fr, hiu), pv = clamp_pad(vn * fr, hiv);
                    float fu = floorf(pu), fv = floorf(pv);
                    float val = bilerp_pairs(ld4b(pl, pair_off(fu, fv, pitchf)), pu - fu, pv - fv);
                    acc = fmaf((KIND == 2) ? wa : fr * fr, val, acc);
                }
            }
        }
    }
    const size_t idx = ((size_t)Z * A.ny + Y) * A.nx + X;
    const float val = acc * A.osc / (float)(R * R * R);
    if (A.mode == OUT_ADD) {
        A.vol[idx] += val;
    } else if (A.mode == OUT_SET) {
        A.vol[idx] = val;
    } else if (A.mode == OUT_EMUPD) {
        A.vol[idx] *= val * A.uw[idx];
    } else {
        float w = A.uw ? A.uw[idx] : 1.f;
        A.vol[idx] = fminf(fmaxf(fmaf(w, val, A.vol[idx]), A.lo), A.hi);
    }
}

// --------------------------------------------------- matched BP ----------
// Exact transpose of fp3d_kernel: every (ray, slice) sample scatters its
// four bilinear weights into the padded slice-major accumulation copy of
// the angle's direction group (fp32 atomics).
template <bool CONE>
__global__ __launch_bounds__(256) void fp3d_transpose_kernel(FPArgs A, const float *__restrict__ proj) {
    unsigned tile;
    int angle;
    if (!fp_map(A, tile, angle)) return;
    const unsigned tiles_u = (A.nu + 15) / 16;
    const unsigned tu = tile % tiles_u, tv = tile / tiles_u;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int u = tu * 16 + (lane & 15);
    const int v = tv * 16 + wave * 4 + (lane >> 4);
    const FP3DAngle &G = A.ang[angle];
    if (u >= (int)A.nu || v >= (int)A.nv) return;
    const float y0 = proj[((size_t)v * A.na + angle) * A.nu + u];
    if (y0 == 0.f) return;
    const unsigned pitch = A.n1 + 2 * kPad;
    const size_t plane = (size_t)(A.n2 + 2 * kPad) * pitch;
    const float pitchf = (float)pitch;
    const float hi1 = (float)A.n1 + 2.5f, hi2 = (float)A.n2 + 2.5f;
    float *acc = const_cast<float *>(A.volp);
    const int R = (int)A.ss;
    for (int k = 0; k < R; ++k)
        for (int l = 0; l < R; ++l) {
            float a1f, a2f, c1f, c2f, scale;
            int lo, hi;
            if (R == 1)
                fp3d_ray<CONE>(A, G, u, v, a1f, a2f, c1f, c2f, scale, lo, hi);
            else
                fp3d_ray<CONE>(A, G, u - 0.5 + (k + 0.5) / R, v - 0.5 + (l + 0.5) / R, a1f, a2f,
                               c1f, c2f, scale, lo, hi);
            const float y = y0 * scale / (float)(R * R);
            for (int s = lo; s <= hi; ++s) {
                const float sf = (float)s;
                float p1 = clamp_pad(fmaf(a1f, sf, c1f), hi1);
                float p2 = clamp_pad(fmaf(a2f, sf, c2f), hi2);
                float f1 = floorf(p1), f2 = floorf(p2);
                float w1 = p1 - f1, w2 = p2 - f2;
                float *q = acc + (size_t)s * plane + (texel_off(f1, f2, pitchf) >> 2);
                atomicAdd(q, y * (1.f - w1) * (1.f - w2));
                atomicAdd(q + 1, y * w1 * (1.f - w2));
                atomicAdd(q + pitch, y * (1.f - w1) * w2);
                atomicAdd(q + pitch + 1, y * w1 * w2);
            }
        }
}

// vol[z][y][x] (+)= interior of a padded slice-major group copy
__global__ __launch_bounds__(256) void unpad_group_add(const float *__restrict__ src,
                                                       float *__restrict__ vol, unsigned nx,
                                                       unsigned ny, unsigned nz, int group) {
    size_t n = (size_t)nx * ny * nz;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        unsigned x = i % nx;
        unsigned y = (i / nx) % ny;
        unsigned z = i / ((size_t)nx * ny);
        size_t s;
        if (group == 0)
            s = ((size_t)x * (nz + 2 * kPad) + z + kPad) * (ny + 2 * kPad) + y + kPad;
        else if (group == 1)
            s = ((size_t)y * (nz + 2 * kPad) + z + kPad) * (nx + 2 * kPad) + x + kPad;
        else
            s = ((size_t)z * (ny + 2 * kPad) + y + kPad) * (nx + 2 * kPad) + x + kPad;
        vol[i] += src[s];
    }
}

// ------------------------------------------------------------ launch -----
namespace {

// dims of direction group g along (c0, c1, c2)
void group_dims(const Geom3D &g, int grp, unsigned &n0, unsigned &n1, unsigned &n2) {
    if (grp == 0) { n0 = g.nx; n1 = g.ny; n2 = g.nz; }
    else if (grp == 1) { n0 = g.ny; n1 = g.nx; n2 = g.nz; }
    else { n0 = g.nz; n1 = g.nx; n2 = g.ny; }
}

size_t group_elems(const Geom3D &g, int grp) {
    unsigned n0, n1, n2;
    group_dims(g, grp, n0, n1, n2);
    return (size_t)n0 * (n1 + 2 * kPad) * (n2 + 2 * kPad);
}

void make_group_copy(const Geom3D &g, int grp, const float *vol, float2 *dst, hipStream_t s) {
    if (grp == 0) {
        dim3 grid((g.ny + 2 * kPad + 63) / 64, (g.nx + 63) / 64, g.nz + 2 * kPad);
        pad_transpose_x<<<grid, 256, 0, s>>>(vol, dst, g.nx, g.ny, g.nz);
    } else {
        unsigned rows = (grp == 2) ? g.nz * (g.ny + 2 * kPad) : g.ny * (g.nz + 2 * kPad);
        pad_copy_xfast<<<rows, 256, 0, s>>>(vol, dst, g.nx, g.ny, g.nz, grp);
    }
    AMX_LAUNCHED();
}

// Cube test and per-direction scale factors (cuda/3d/cone_fp.cu:339-370,
// cuda/3d/par3d_fp.cu:438-472).
void fp_scales(const Geom3D &g, int grp, float out_scale, float &fs1, float &fs2, float &osc) {
    const float vx = g.vol_scale[0], vy = g.vol_scale[1], vz = g.vol_scale[2];
    bool cube = !(std::fabs(vx / vy - 1.0) > 0.00001) && !(std::fabs(vx / vz - 1.0) > 0.00001);
    if (cube) {
        fs1 = fs2 = 1.f;
        osc = out_scale * vx;
        return;
    }
    float s1, s2;
    if (grp == 0) {
        s1 = vy / vx; s2 = vz / vx; osc = out_scale * vx;
    } else if (grp == 1) {
        s1 = vx / vy;
        // par3d keeps volScale.fY / volScale.fY here (cuda/3d/par3d_fp.cu:459)
        s2 = (g.kind == ASTRA_AMD_GEOM_PAR3D) ? vy / vy : vz / vy;
        osc = out_scale * vy;
    } else {
        s1 = vx / vz; s2 = vy / vz; osc = out_scale * vz;
    }
    fs1 = s1 * s1;
    fs2 = s2 * s2;
}

// ASTRA_AMD_BP3D_GENERAL=1 forces the general (z-dependent) BP form even
// where the z-constant form applies (for comparing the two).
bool zc_disabled() {
    static const bool v = [] {
        const char *e = std::getenv("ASTRA_AMD_BP3D_GENERAL");
        return e && e[0] == '1';
    }();
    return v;
}

// Angles per block of the FP workgroup map (fp_map); ASTRA_AMD_FP3D_ANGLE_BLOCK
// overrides the default (1..64).
unsigned fp_angle_block() {
    static const unsigned v = [] {
        const char *e = std::getenv("ASTRA_AMD_FP3D_ANGLE_BLOCK");
        long x = e ? std::strtol(e, nullptr, 10) : 0;
        return (x >= 1 && x <= 64) ? (unsigned)x : 16u;
    }();
    return v;
}

// Launch `kern` over the angles of one direction group, in as many
// launches as the 2^32 work-item limit of one grid requires.
template <typename K>
void fp_group_launches(FPArgs A, unsigned count, K kern, hipStream_t s) {
    A.tiles = ((A.nu + 15) / 16) * ((A.nv + 15) / 16);
    A.ablk = std::min(fp_angle_block(), count);
    const size_t per_block = (size_t)A.tiles * A.ablk;
    const size_t max_groups = (1u << 24) - 8;  // 256 threads each
    AMX_REQUIRE(per_block <= max_groups, "fp3d: detector too large for one launch");
    const unsigned blocks_per_launch = (unsigned)(max_groups / per_block);
    for (unsigned first = 0; first < count; first += blocks_per_launch * A.ablk) {
        A.lfirst = first;
        A.lcount = std::min(count - first, blocks_per_launch * A.ablk);
        const unsigned nblk = (A.lcount + A.ablk - 1) / A.ablk;
        const unsigned groups = (unsigned)(((size_t)nblk * per_block + 7) & ~(size_t)7);
        kern(A, groups);
        AMX_LAUNCHED();
    }
}

FPArgs fp_args(const Geom3D &g, int grp, float out_scale) {
    FPArgs A{};
    group_dims(g, grp, A.n0, A.n1, A.n2);
    A.ang = g.fp.as<FP3DAngle>();
    A.alist = g.group_idx[grp].as<int>();
    A.nu = g.nu; A.nv = g.nv; A.na = g.na;
    fp_scales(g, grp, out_scale, A.fs1, A.fs2, A.osc);
    return A;
}

}  // namespace

void fp3d_launch(const Geom3D &g, const float *vol, float *proj, float out_scale, int mode,
                 const float *resid_b, const float *resid_w, bool padded_out, hipStream_t s,
                 unsigned ss) {
    AMX_REQUIRE(ss >= 1 && ss <= 16, "fp3d: rays per detector dimension must be in [1, 16]");
    size_t need = 0;
    int groups[3], ng = 0;
    for (int k = 0; k < 3; ++k)
        if (g.dir_count[k]) {
            need = std::max(need, group_elems(g, k));
            groups[ng++] = k;
        }
    Scratch copy(need * sizeof(float2), s);  // row-pair copy
    for (int i = 0; i < ng; ++i) {
        const int k = groups[i];
        make_group_copy(g, k, vol, copy.as<float2>(), s);
        FPArgs A = fp_args(g, k, out_scale);
        A.volp = copy.as<float>();
        A.out = proj;
        A.rb = resid_b;
        A.rw = resid_w;
        A.padded_out = padded_out ? 1 : 0;
        // The direction groups partition the angles: every output element
        // is written by exactly one group, which applies the full mode.
        A.mode = mode;
        ProfToken pt = prof_begin(g.kind == ASTRA_AMD_GEOM_CONE ? "fp3d_cone" : "fp3d_par3d",
                                  (double)g.dir_count[k] * g.nu * g.nv * A.n0 * ss * ss, s);
        const bool cone = g.kind == ASTRA_AMD_GEOM_CONE;
        A.ss = ss;
        fp_group_launches(A, g.dir_count[k], [&](const FPArgs &a, unsigned groups) {
            if (ss > 1) {
                if (cone)
                    fp3d_kernel<true, true><<<groups, 256, 0, s>>>(a);
                else
                    fp3d_kernel<false, true><<<groups, 256, 0, s>>>(a);
            } else {
                if (cone)
                    fp3d_kernel<true, false><<<groups, 256, 0, s>>>(a);
                else
                    fp3d_kernel<false, false><<<groups, 256, 0, s>>>(a);
            }
        }, s);
        prof_end(pt, s);
    }
}

void fp3d_sumsq_launch(const Geom3D &g, float *proj, float out_scale, int mode, hipStream_t s) {
    AMX_REQUIRE(g.kind == ASTRA_AMD_GEOM_PAR3D,
                "fp3d: sum_square_weights is defined for parallel3d geometries only");
    AMX_REQUIRE(mode == OUT_SET || mode == OUT_ADD, "fp3d: sum_square_weights mode");
    for (int k = 0; k < 3; ++k) {
        if (!g.dir_count[k]) continue;
        FPArgs A = fp_args(g, k, out_scale);
        A.out = proj;
        A.mode = mode;
        A.ss = 1;
        fp_group_launches(A, g.dir_count[k], [&](const FPArgs &a, unsigned groups) {
            fp3d_sumsq_kernel<<<groups, 256, 0, s>>>(a);
        }, s);
    }
}

void proj_to_padded_am(const Geom3D &g, const float *proj, float *pp, hipStream_t s) {
    const unsigned pv = g.nv + 2 * kPad;
    proj_to_am<<<g.na * pv, 256, 0, s>>>(proj, pp, g.nu, g.na, g.nv);
    AMX_LAUNCHED();
}

size_t padded_am_elems(const Geom3D &g) {
    return (size_t)g.na * (g.nv + 2 * kPad) * (g.nu + 2 * kPad);
}

namespace {
// BP from a row-pair angle-major copy
void bp3d_run_pairs(const Geom3D &g, const float2 *pairs, float *vol, float out_scale, int kernel,
                    int mode, const float *upd_w, float lo, float hi, hipStream_t s, unsigned ss) {
    AMX_REQUIRE(ss >= 1 && ss <= 16, "bp3d: rays per voxel dimension must be in [1, 16]");
    BPArgs A{};
    A.pp = reinterpret_cast<const float *>(pairs);
    A.nu = g.nu; A.nv = g.nv; A.na = g.na;
    A.nx = g.nx; A.ny = g.ny; A.nz = g.nz;
    A.vol = vol;
    A.mode = mode;
    A.uw = upd_w;
    A.lo = lo;
    A.hi = hi;
    const float vx = g.vol_scale[0], vy = g.vol_scale[1], vz = g.vol_scale[2];
    // ConeBP_Array / Par3DBP_Array output scale (cuda/3d/cone_bp.cu:322-328,
    // cuda/3d/par3d_bp.cu:263)
    if (g.kind == ASTRA_AMD_GEOM_CONE && kernel == ASTRA_AMD_KER3D_FDK_WEIGHTING)
        A.osc = out_scale / vx;
    else
        A.osc = out_scale * (vx * vy * vz);
    constexpr int ZS = 8;
    dim3 grid(((g.nx + 15) / 16) * ((g.ny + 15) / 16), (g.nz + ZS - 1) / ZS);
    const char *pname = g.kind != ASTRA_AMD_GEOM_CONE ? "bp3d_par3d"
                        : (kernel == ASTRA_AMD_KER3D_FDK_WEIGHTING ? "bp3d_fdk" : "bp3d_cone");
    ProfToken pt = prof_begin(pname, (double)g.nx * g.ny * g.nz * g.na, s);
    int kind;
    bool zc;
    if (g.kind == ASTRA_AMD_GEOM_CONE) {
        if (kernel == ASTRA_AMD_KER3D_FDK_WEIGHTING) {
            A.ang = g.bp_fdk.as<BP3DAngle>();
            kind = 0;
            zc = g.zc_fdk;
        } else if (kernel == ASTRA_AMD_KER3D_2D_WEIGHTING) {
            A.ang = g.bp_2d.as<BP3DAngle>();
            kind = 1;
            zc = g.zc_2d;
        } else {
            A.ang = g.bp_default.as<BP3DAngle>();
            kind = 0;
            zc = g.zc_default;
        }
    } else {
        if (kernel == ASTRA_AMD_KER3D_2D_WEIGHTING) {
            A.ang = g.bp_2d.as<BP3DAngle>();
            A.aw = g.bp_w_2d.as<float>();
            zc = g.zc_2d;
        } else {
            A.ang = g.bp_default.as<BP3DAngle>();
            A.aw = g.bp_w_default.as<float>();
            zc = g.zc_default;
        }
        kind = 2;
    }
    if (zc_disabled()) zc = false;
    // supersampling applies to the default cone kernel and to par3d, as in
    // ConeBP_Array / Par3DBP_Array (cuda/3d/cone_bp.cu:351-370,
    // cuda/3d/par3d_bp.cu:285-293)
    const bool use_ss = ss > 1 && (g.kind != ASTRA_AMD_GEOM_CONE || kernel == ASTRA_AMD_KER3D_DEFAULT);
    if (use_ss) {
        dim3 sgrid(((g.nx + 15) / 16) * ((g.ny + 15) / 16), g.nz);
        if (kind == 0)
            bp3d_ss_kernel<0><<<sgrid, 256, 0, s>>>(A, (int)ss);
        else
            bp3d_ss_kernel<2><<<sgrid, 256, 0, s>>>(A, (int)ss);
    } else if (kind == 0)
        zc ? bp3d_kernel<0, ZS, true><<<grid, 256, 0, s>>>(A) : bp3d_kernel<0, ZS, false><<<grid, 256, 0, s>>>(A);
    else if (kind == 1)
        zc ? bp3d_kernel<1, ZS, true><<<grid, 256, 0, s>>>(A) : bp3d_kernel<1, ZS, false><<<grid, 256, 0, s>>>(A);
    else
        zc ? bp3d_kernel<2, ZS, true><<<grid, 256, 0, s>>>(A) : bp3d_kernel<2, ZS, false><<<grid, 256, 0, s>>>(A);
    AMX_LAUNCHED();
    prof_end(pt, s);
}

}  // namespace

void bp3d_launch_padded(const Geom3D &g, const float *pp, float *vol, float out_scale, int kernel,
                        int mode, const float *upd_w, float lo, float hi, hipStream_t s,
                        unsigned ss) {
    Scratch pairs(padded_am_elems(g) * sizeof(float2), s);
    am_to_pairs<<<g.na * (g.nv + 2 * kPad), 256, 0, s>>>(pp, pairs.as<float2>(), g.nu, g.nv);
    AMX_LAUNCHED();
    bp3d_run_pairs(g, pairs.as<float2>(), vol, out_scale, kernel, mode, upd_w, lo, hi, s, ss);
}

void bp3d_launch(const Geom3D &g, const float *proj, float *vol, float out_scale, int kernel,
                 int mode, const float *upd_w, float lo, float hi, hipStream_t s, unsigned ss) {
    Scratch pairs(padded_am_elems(g) * sizeof(float2), s);
    proj_to_am_pairs<<<g.na * (g.nv + 2 * kPad), 256, 0, s>>>(proj, pairs.as<float2>(), g.nu, g.na,
                                                               g.nv);
    AMX_LAUNCHED();
    bp3d_run_pairs(g, pairs.as<float2>(), vol, out_scale, kernel, mode, upd_w, lo, hi, s, ss);
}

void bp3d_matched_launch(const Geom3D &g, const float *proj, float *vol, float out_scale,
                         int mode, hipStream_t s, unsigned ss) {
    AMX_REQUIRE(ss >= 1 && ss <= 16, "bp3d: rays per detector dimension must be in [1, 16]");
    const size_t nvox = (size_t)g.nx * g.ny * g.nz;
    if (mode == OUT_SET) fill_launch(vol, nvox, 0.f, s);
    size_t need = 0;
    for (int k = 0; k < 3; ++k)
        if (g.dir_count[k]) need = std::max(need, group_elems(g, k));
    Scratch acc(need * sizeof(float), s);
    for (int k = 0; k < 3; ++k) {
        if (!g.dir_count[k]) continue;
        fill_launch(acc.as<float>(), group_elems(g, k), 0.f, s);
        FPArgs A = fp_args(g, k, out_scale);
        A.volp = acc.as<float>();
        A.ss = ss;
        const bool cone = g.kind == ASTRA_AMD_GEOM_CONE;
        fp_group_launches(A, g.dir_count[k], [&](const FPArgs &a, unsigned groups) {
            if (cone)
                fp3d_transpose_kernel<true><<<groups, 256, 0, s>>>(a, proj);
            else
                fp3d_transpose_kernel<false><<<groups, 256, 0, s>>>(a, proj);
        }, s);
        unsigned blocks = (unsigned)std::min<size_t>((nvox + 255) / 256, 1u << 20);
        unpad_group_add<<<blocks, 256, 0, s>>>(acc.as<float>(), vol, g.nx, g.ny, g.nz, k);
        AMX_LAUNCHED();
    }
}

}  // namespace amx
