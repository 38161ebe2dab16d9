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
//   Both copies are stored as row difference pairs: element [r][c] is the
//   float2 (p[r][c], p[r+1][c] - p[r][c]), so the four taps of a bilinear
//   sample at (r, c) are one 16-byte load of [r][c] and [r][c+1] (the
//   texture-address unit handles one instruction per sample instead of
//   two), and each row-direction lerp is one fma; the copies take twice
//   the memory.
//
// Sampling semantics follow cuda/3d/cone_fp.cu:154-231,
// cuda/3d/par3d_fp.cu:152-230, cuda/3d/cone_bp.cu:55-133 and
// cuda/3d/par3d_bp.cu:55-120 with ideal fp32 (bi)linear interpolation in
// place of the 8-bit-weight texture unit.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <memory>

#include "devutil.hpp"
#include "internal.hpp"

namespace amx {

// ----------------------------------------------------------- prep --------
// dst[c0][c2 + 2][c1 + 2] = (vol at c2, vol at c2 + 1 - vol at c2) for the
// z- and y-groups (the fast axis stays x; difference pairs as described
// above).  One workgroup per destination row; borders read 0.
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
        const float a = (in0 && xin) ? s0[x] : 0.f, b = (in1 && xin) ? s1[x] : 0.f;
        out[i] = make_float2(a, b - a);
    }
}

// x-group: dst[x][z + 2][y + 2] = (vol[z][y][x], vol[z+1][y][x] - vol[z][y][x]); an
// (x, y) transpose of a 64 x 64 tile for kTzRows consecutive padded z rows,
// each volume plane of the tile read once into one of two LDS tiles (row
// zr combines planes zr - 2 and zr - 1).  grid = (y tiles, x tiles,
// ceil((nz + 4) / kTzRows)).
constexpr int kTzRows = 8;
__global__ __launch_bounds__(256) void pad_transpose_x(const float *__restrict__ vol,
                                                       float2 *__restrict__ dst, unsigned nx,
                                                       unsigned ny, unsigned nz) {
    __shared__ float tile[2][64][65];
    const unsigned pyn = ny + 2 * kPad, pzn = nz + 2 * kPad;
    const int zr0 = blockIdx.z * kTzRows;  // first padded z row
    const int x0 = blockIdx.y * 64;
    const int yp0 = blockIdx.x * 64;  // padded y tile origin
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int nrows = min(kTzRows, (int)pzn - zr0);
    // plane k (k = 0 .. nrows) is z = zr0 - kPad + k, held in tile[k & 1]
    for (int k = 0; k <= nrows; ++k) {
        const int z = zr0 - kPad + k;
        float *t = &tile[k & 1][0][0];
        for (int r = ty; r < 64; r += 4) {
            const int y = yp0 + r - kPad, x = x0 + tx;
            float v = 0.f;
            if (z >= 0 && z < (int)nz && y >= 0 && y < (int)ny && x < (int)nx)
                v = vol[((size_t)z * ny + y) * nx + x];
            t[r * 65 + tx] = v;
        }
        __syncthreads();
        if (k > 0) {
            // padded row zr0 + k - 1 from planes k - 1 and k
            const int zr = zr0 + k - 1;
            const float *a = &tile[(k - 1) & 1][0][0], *b = &tile[k & 1][0][0];
            for (int r = ty; r < 64; r += 4) {
                const int x = x0 + r, yp = yp0 + tx;
                if (x < (int)nx && yp < (int)pyn) {
                    const float va = a[tx * 65 + r];
                    dst[((size_t)x * pzn + zr) * pyn + yp] = make_float2(va, b[tx * 65 + r] - va);
                }
            }
        }
        __syncthreads();  // tile[(k + 1) & 1] is overwritten next
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

// proj [v][a][u] -> padded angle-major difference pairs [a][v + 2][u + 2] =
// (p[v][u], p[v+1][u] - p[v][u]) (borders read 0).  One workgroup per
// destination row.
// A workgroup writes kPairRows consecutive padded rows of one angle: each
// thread walks a column strip down the rows, so every source row is read
// once (plus one row per strip) and the row loads are independent.
// grid = na * ceil((nv + 4) / kPairRows).
constexpr int kPairRows = 8;
__global__ __launch_bounds__(256) void proj_to_am_pairs(const float *__restrict__ proj,
                                                        float2 *__restrict__ pp, unsigned nu,
                                                        unsigned na, unsigned nv) {
    const unsigned pu = nu + 2 * kPad, pv = nv + 2 * kPad;
    const unsigned blocks = (pv + kPairRows - 1) / kPairRows;
    const unsigned a = blockIdx.x / blocks;
    const int r0 = (int)(blockIdx.x % blocks) * kPairRows;  // first padded row
    const int nrows = min(kPairRows, (int)pv - r0);
    float2 *out = pp + ((size_t)a * pv + r0) * pu;
    // source row of padded row r (zero outside the data)
    auto src = [&](int r) -> const float * {
        const int v = r - kPad;
        return (v >= 0 && v < (int)nv) ? proj + ((size_t)v * na + a) * nu : nullptr;
    };
    for (unsigned i = threadIdx.x; i < pu; i += blockDim.x) {
        const int u = (int)i - kPad;
        const bool uin = u >= 0 && u < (int)nu;
        float cur[kPairRows + 1];
#pragma unroll
        for (int k = 0; k <= kPairRows; ++k) {
            const float *s = k <= nrows ? src(r0 + k) : nullptr;
            cur[k] = (s && uin) ? s[u] : 0.f;
        }
#pragma unroll
        for (int k = 0; k < kPairRows; ++k)
            if (k < nrows) out[(size_t)k * pu + i] = make_float2(cur[k], cur[k + 1] - cur[k]);
    }
}

// padded single layout [a][v + 4][u + 4] -> difference pairs (same padding);
// kPairRows rows of one angle per workgroup, as proj_to_am_pairs.
// grid = na * ceil((nv + 4) / kPairRows).
__global__ __launch_bounds__(256) void am_to_pairs(const float *__restrict__ src,
                                                   float2 *__restrict__ pp, unsigned nu,
                                                   unsigned nv) {
    const unsigned pu = nu + 2 * kPad, pv = nv + 2 * kPad;
    const unsigned blocks = (pv + kPairRows - 1) / kPairRows;
    const unsigned a = blockIdx.x / blocks;
    const int r0 = (int)(blockIdx.x % blocks) * kPairRows;  // first padded row
    const int nrows = min(kPairRows, (int)pv - r0);
    const float *s0 = src + ((size_t)a * pv + r0) * pu;
    float2 *out = pp + ((size_t)a * pv + r0) * pu;
    for (unsigned i = threadIdx.x; i < pu; i += blockDim.x) {
        float cur[kPairRows + 1];
#pragma unroll
        for (int k = 0; k <= kPairRows; ++k)  // row pv (past the last) reads 0
            cur[k] = (k <= nrows && r0 + k < (int)pv) ? s0[(size_t)k * pu + i] : 0.f;
#pragma unroll
        for (int k = 0; k < kPairRows; ++k)
            if (k < nrows) out[(size_t)k * pu + i] = make_float2(cur[k], cur[k + 1] - cur[k]);
    }
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
    // workgroup map (fp_map): detector tiles, angles per block, the range
    // [lfirst, lfirst + lcount) of alist covered by this launch, and the
    // part [xfirst[x], xfirst[x + 1]) of that range taken by XCD x
    unsigned tiles, ablk, lfirst, lcount;
    unsigned xfirst[9];
    unsigned ss;  // rays per detector pixel and dimension (DetectorSuperSampling)
};

// Workgroup -> (detector tile, angle) map.  Workgroups are handed out
// round-robin over the 8 XCDs (workgroup L runs on XCD L % 8), and each XCD
// has its own L2.  XCD x takes the angles [xfirst[x], xfirst[x + 1]) of the
// launch's range (an eighth of them), cut into equal blocks of at most
// `ablk` angles, in the order (angle block, tile, angle within block); no
// block spans two XCDs.  The workgroups resident on one XCD at a time are
// then a few neighbouring tiles, each over a block of consecutive angles,
// which start together and march through the same slices at the same
// time: rays of neighbouring angles sample nearly the same voxels, so they
// share the XCD's L2 instead of each fetching the volume from HBM.
// Workgroups beyond their XCD's range return at once.  (A single range cut
// into blocks across the XCDs left the last XCD with the partial block and
// the others with 48 of every 45 angles' work: 48.7 vs 42.7 ms per group
// launch at 512^3 / 720 x 768^2.)
__device__ __forceinline__ bool fp_map_L(const FPArgs &A, unsigned L, unsigned &tile, int &angle) {
    const unsigned x = L & 7u, q = L >> 3;
    const unsigned first = A.xfirst[x], count = A.xfirst[x + 1] - first;
    if (count == 0) return false;
    const unsigned nblk = (count + A.ablk - 1) / A.ablk;
    const unsigned bsz = (count + nblk - 1) / nblk;  // angles per block
    const unsigned span = bsz * A.tiles;             // items per full block
    const unsigned blk = q / span;
    const unsigned b0 = blk * bsz;
    if (b0 >= count) return false;
    const unsigned bs = min(bsz, count - b0);
    const unsigned r = q - blk * span;
    if (r >= bs * A.tiles) return false;
    tile = r / bs;
    angle = A.alist[A.lfirst + first + b0 + r % bs];
    return true;
}

__device__ __forceinline__ bool fp_map(const FPArgs &A, unsigned &tile, int &angle) {
    return fp_map_L(A, blockIdx.x, tile, angle);
}

// One thread per ray; a 256-thread workgroup covers a 16 x 16 detector tile
// of one angle, each wave a 16 u x 4 v sub-tile.  Ray setup runs in FP64;
// the slice loop is wave-uniform over the union of the lanes' clipped slice
// ranges, so the slice base address is scalar and every fetch is one
// 16-byte load with a 32-bit vector offset.
// (u, v): detector coordinates of the ray relative to pixel (0, 0)'s centre
// (fractional for sub-rays).
// GK: GEO_PAR3D, GEO_CONE (flat detector) or GEO_CYL (cylindrical detector:
// pixel centre on the cylinder, cuda/3d/cone_cyl.cu:243-258).
enum { GEO_PAR3D = 0, GEO_CONE = 1, GEO_CYL = 2 };

// The ray of detector coordinates (u, v) in the group's slice frame:
// texel coordinate - 0.5 along axes 1, 2 at slice s = a_k s + c_k, and the
// slices [lo, hi] where it can touch the volume.
template <int GK>
__device__ __forceinline__ void fp3d_ray_line(const FP3DAngle &G, unsigned nu, unsigned nv,
                                              unsigned n0, unsigned n1, unsigned n2, double u,
                                              double v, double &a1, double &a2, double &c1,
                                              double &c2, int &lo, int &hi) {
    constexpr bool CONE = GK != GEO_PAR3D;
    double P[3], d[3];
    if (GK == GEO_CYL) {
        const double th = (u - 0.5 * nu + 0.5) * G.dth;
        const double fv = v - 0.5 * nv + 0.5;
        const double ct = cos(th), st = sin(th);
        for (int k = 0; k < 3; ++k) P[k] = G.p0[k] + ct * G.u[k] + st * G.b[k] + fv * G.v[k];
    } else {
        for (int k = 0; k < 3; ++k) P[k] = G.p0[k] + u * G.u[k] + v * G.v[k];
    }
    for (int k = 0; k < 3; ++k) d[k] = CONE ? (G.s[k] - P[k]) : G.s[k];
    const double *B = CONE ? G.s : P;
    a1 = d[1] / d[0];
    a2 = d[2] / d[0];
    // texel coordinate - 0.5 at slice s: a_k (s - n0/2 + 0.5 - B0) + B_k + n_k/2 - 0.5
    const double base = -0.5 * n0 + 0.5 - B[0];
    c1 = a1 * base + B[1] + 0.5 * n1 - 0.5;
    c2 = a2 * base + B[2] + 0.5 * n2 - 0.5;
    lo = 0;
    hi = (int)n0 - 1;
    clip_range(a1, c1, (double)n1, (int)n0, lo, hi);
    clip_range(a2, c2, (double)n2, (int)n0, lo, hi);
}

template <int GK>
__device__ __forceinline__ void fp3d_ray(const FPArgs &A, const FP3DAngle &G, double u, double v,
                                         float &a1f, float &a2f, float &c1f, float &c2f,
                                         float &scale, int &lo, int &hi) {
    double a1, a2, c1, c2;
    fp3d_ray_line<GK>(G, A.nu, A.nv, A.n0, A.n1, A.n2, u, v, a1, a2, c1, c2, lo, hi);
    a1f = (float)a1;
    a2f = (float)a2;
    c1f = (float)(c1 + kPad);  // padded texel coordinates
    c2f = (float)(c2 + kPad);
    scale = (float)(sqrt(a1 * a1 * A.fs1 + a2 * a2 * A.fs2 + 1.0) * A.osc);
}

// Sum of the bilinear samples of slices [slo, shi] (wave-uniform) along the
// ray (c1, c2) = a s + c through the padded group copy.  Samples at even
// offsets from slo go to acc0, odd ones to acc1; four loads are issued
// before the first of them is used (the loop is bound by the loads'
// latency).  The loads are structured buffer loads of the pair records:
// the slice base and the slice offsets k * plane are scalar, the record
// index (one v_mad_u32_u24) the only per-lane address term.
__device__ __forceinline__ float march3d(const FPArgs &A, float a1f, float a2f, float c1f,
                                         float c2f, int slo, int shi) {
    const unsigned pitch = A.n1 + 2 * kPad;                        // pairs per row
    const size_t plane = (size_t)(A.n2 + 2 * kPad) * pitch * 2;     // floats per slice
    // bytes per slice (< 1 GiB: host check), wave-uniform
    const unsigned plane_b = __builtin_amdgcn_readfirstlane((unsigned)(plane * sizeof(float)));
    const float hi1 = (float)A.n1 + 2.5f, hi2 = (float)A.n2 + 2.5f;
    float acc0 = 0.f, acc1 = 0.f;
    int s = slo;
    const char *pl =
        first_lane_ptr(reinterpret_cast<const char *>(A.volp + (size_t)(slo > 0 ? slo : 0) * plane));
    for (; s + 3 <= shi; s += 4) {
        const __amdgpu_buffer_rsrc_t rs = pair_rsrc(pl);
        float4 q[4];
        float w1[4], w2[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float sf = (float)(s + k);
            const float p1 = clamp_pad(fmaf(a1f, sf, c1f), hi1);
            const float p2 = clamp_pad(fmaf(a2f, sf, c2f), hi2);
            w1[k] = frac(p1);
            w2[k] = frac(p2);
            q[k] = ld4p(rs, pair_idx(p1, p2, pitch), k * plane_b);
        }
        acc0 += bilerp_pairs(q[0], w1[0], w2[0]);
        acc1 += bilerp_pairs(q[1], w1[1], w2[1]);
        acc0 += bilerp_pairs(q[2], w1[2], w2[2]);
        acc1 += bilerp_pairs(q[3], w1[3], w2[3]);
        pl += 4 * (size_t)plane_b;
    }
    // the last one to three slices
    const __amdgpu_buffer_rsrc_t rs = pair_rsrc(pl);
    unsigned so = 0;
    if (s + 1 <= shi) {
        const float sf = (float)s;
        float p1 = clamp_pad(fmaf(a1f, sf, c1f), hi1);
        float p2 = clamp_pad(fmaf(a2f, sf, c2f), hi2);
        float q1 = clamp_pad(fmaf(a1f, sf + 1.f, c1f), hi1);
        float q2 = clamp_pad(fmaf(a2f, sf + 1.f, c2f), hi2);
        float4 r = ld4p(rs, pair_idx(p1, p2, pitch), 0u);
        float4 t = ld4p(rs, pair_idx(q1, q2, pitch), plane_b);
        acc0 += bilerp_pairs(r, frac(p1), frac(p2));
        acc1 += bilerp_pairs(t, frac(q1), frac(q2));
        so = 2 * plane_b;
        s += 2;
    }
    if (s <= shi) {
        const float sf = (float)s;
        float p1 = clamp_pad(fmaf(a1f, sf, c1f), hi1);
        float p2 = clamp_pad(fmaf(a2f, sf, c2f), hi2);
        acc0 += bilerp_pairs(ld4p(rs, pair_idx(p1, p2, pitch), so), frac(p1), frac(p2));
    }
    return acc0 + acc1;
}

// SS: A.ss x A.ss sub-rays per detector pixel at offsets (k + 0.5) / ss - 0.5
// in u and v, each with its own distance weight, averaged
// (cone_FP_SS_t / par3D_FP_SS_t, cuda/3d/cone_fp.cu:233-323,
// cuda/3d/par3d_fp.cu:234-325).
template <int GK, bool SS>
__device__ __forceinline__ void fp3d_item(const FPArgs &A, unsigned L) {
    unsigned tile;
    int angle;
    if (!fp_map_L(A, L, tile, angle)) return;  // uniform over the workgroup
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
        if (active) fp3d_ray<GK>(A, G, u, v, a1f, a2f, c1f, c2f, scale, lo, hi);
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
                    fp3d_ray<GK>(A, G, u - 0.5 + (k + 0.5) / R, v - 0.5 + (l + 0.5) / R, a1f,
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

// The FP as a resident grid: workgroup b starts on logical items of XCD
// b % 8 (the items fp_map gives workgroup L = x + 8 t on XCD x), taking
// t from that XCD's ticket counter, and when its XCD's items are taken it
// takes the next XCD's.  Items do not depend on each other, so no
// workgroup ever waits for another; every workgroup leaves once all
// tickets are taken.  The same rays and sums as one workgroup per item
// (bit for bit), with no XCD left idle while another still has items
// (41.8 vs 43.0 ms per group launch at 512^3 / 720 x 768^2).
template <int GK, bool SS>
__global__ __launch_bounds__(256) void fp3d_kernel(FPArgs A, unsigned *tick, unsigned per_xcd) {
    __shared__ unsigned sh;
    const unsigned home = blockIdx.x & 7u;
    unsigned q = 0;  // XCD ticket counters exhausted so far (thread 0)
    for (;;) {
        if (threadIdx.x == 0) {
            unsigned L = 0xffffffffu;
            while (q < 8) {
                const unsigned x = (home + q) & 7u;
                const unsigned t = atomicAdd(tick + x, 1u);
                if (t < per_xcd) {
                    L = x + 8u * t;
                    break;
                }
                ++q;
            }
            sh = L;
        }
        __syncthreads();
        const unsigned L = __builtin_amdgcn_readfirstlane(sh);
        if (L == 0xffffffffu) break;
        fp3d_item<GK, SS>(A, L);
        __syncthreads();  // every thread has read sh
    }
}

__global__ void fp_tick_reset(unsigned *tick) {
    if (threadIdx.x < 8) tick[threadIdx.x] = 0;
}

// ---- FP with volume slices staged in LDS ---------------------------------
// A workgroup takes a 16 x 16 detector tile at kFpAngles consecutive angles
// of its direction group (each thread: one detector pixel at every one of
// those angles; each wave: the 16 u x 4 v rays of fp3d_kernel's waves).
// Neighbouring angles sample nearly the same voxels, so the workgroup
// marches the slices in lockstep, kFpSlices at a time: for every slice of
// a chunk it stages the rectangle of 16-byte quads (pair (r, c) and pair
// (r, c + 1), one whole bilinear tap) that its rays can sample -- the
// bounding box of the clamped sample positions of the corner rays of every
// angle (within a slice the sample position is linear-fractional in the
// detector position, so its extremes over the tile are at the corners),
// widened by one texel -- and every tap is one aligned 16-byte LDS read.
// The next chunk's quads are loaded into registers while the current one
// is sampled.  Each ray sums exactly the samples of march3d in the same
// order (the wave's slice range, alternating accumulators), so the values
// equal fp3d_kernel's.  A chunk whose rectangle exceeds the tile samples
// global memory.
constexpr int kFpAngles = 4;          // angles per workgroup
constexpr int kFpSlices = 4;          // slices per staged chunk
constexpr int kFpQuadCap = 1536;      // LDS quads per chunk (24 KiB)
constexpr int kFpSlots = kFpQuadCap / 256;

struct FpBox {
    int u0[kFpSlices], v0[kFpSlices];  // per slice: first quad column, first row
    int w, h, lp;                      // common width, height, LDS row pitch (quads)
};

// workgroup -> (detector tile, first of kFpAngles angles), XCD-aware as
// fp_map: the workgroups resident on one XCD take neighbouring angle
// groups of the same tiles
__device__ __forceinline__ bool fp_lds_map(const FPArgs &A, unsigned &tile, unsigned &li0) {
    const unsigned L = blockIdx.x;
    const unsigned per = gridDim.x >> 3;
    const unsigned logical = (L & 7u) * per + (L >> 3);
    constexpr unsigned sub = 4;  // angle groups per block
    const unsigned a_in = logical % sub;
    const unsigned rest = logical / sub;
    tile = rest % A.tiles;
    li0 = ((rest / A.tiles) * sub + a_in) * kFpAngles;
    return li0 < A.lcount;
}

template <int GK>
__global__ __launch_bounds__(256) void fp3d_lds_kernel(FPArgs A) {
    __shared__ float4 tile[kFpQuadCap];
    __shared__ float cpar[kFpAngles][4][4];  // corner rays: a1, c1, a2, c2
    __shared__ int crng[kFpAngles][4];       // corner rays: 1 if the angle exists
    __shared__ FpBox box[2];
    __shared__ int wr[4][2];                 // per wave: slice range of the group
    unsigned tile_id, li0;
    if (!fp_lds_map(A, tile_id, li0)) return;  // uniform over the workgroup
    const unsigned tid = threadIdx.x;
    const unsigned tiles_u = (A.nu + 15) / 16;
    const unsigned tu = tile_id % tiles_u, tv = tile_id / tiles_u;
    const int lane = tid & 63, wave = tid >> 6;
    const int u = tu * 16 + (lane & 15);
    const int v = tv * 16 + wave * 4 + (lane >> 4);
    const bool pix = u < (int)A.nu && v < (int)A.nv;
    // last pixel of the tile inside the detector (the corner rays)
    const int ulast = min((int)tu * 16 + 15, (int)A.nu - 1), vlast = min((int)tv * 16 + 15, (int)A.nv - 1);
    // corner q of the tile: u = (q & 1 ? ulast : first), v = (q & 2 ? vlast : first)
    unsigned corners = 0;
    for (int q = 0; q < 4; ++q)
        if (pix && u == ((q & 1) ? ulast : (int)tu * 16) && v == ((q & 2) ? vlast : (int)tv * 16))
            corners |= 1u << q;

    float a1[kFpAngles], a2[kFpAngles], c1[kFpAngles], c2[kFpAngles], sc[kFpAngles];
    int slo[kFpAngles], shi[kFpAngles], angle[kFpAngles];
    int glo = (int)A.n0, ghi = -1;
#pragma unroll
    for (int j = 0; j < kFpAngles; ++j) {
        const unsigned li = li0 + j;
        const bool on = li < A.lcount;
        angle[j] = on ? A.alist[A.lfirst + li] : -1;
        a1[j] = a2[j] = 0.f;
        c1[j] = c2[j] = 0.5f;
        sc[j] = 0.f;
        int lo = (int)A.n0, hi = -1;
        if (on && pix) fp3d_ray<GK>(A, A.ang[angle[j]], u, v, a1[j], a2[j], c1[j], c2[j], sc[j], lo, hi);
        slo[j] = __builtin_amdgcn_readfirstlane(wave_min(lo));
        shi[j] = __builtin_amdgcn_readfirstlane(wave_max(hi));
        glo = min(glo, slo[j]);
        ghi = max(ghi, shi[j]);
        // (a tile of one pixel column / row has coinciding corners: one
        // lane then fills several slots)
        for (int q = 0; q < 4; ++q) {
            if (!(corners & (1u << q))) continue;
            cpar[j][q][0] = a1[j];
            cpar[j][q][1] = c1[j];
            cpar[j][q][2] = a2[j];
            cpar[j][q][3] = c2[j];
            // the clamped positions bound every ray of the tile at every
            // slice, whether or not the corner ray itself meets the volume
            crng[j][q] = on ? 1 : 0;
        }
    }
    if (lane == 0) {
        wr[wave][0] = glo;
        wr[wave][1] = ghi;
    }
    __syncthreads();
    glo = min(min(wr[0][0], wr[1][0]), min(wr[2][0], wr[3][0]));
    ghi = max(max(wr[0][1], wr[1][1]), max(wr[2][1], wr[3][1]));
    if (glo > ghi) return;  // no ray of the group meets the volume (uniform)

    const unsigned pitch = A.n1 + 2 * kPad;                  // pairs per row
    const unsigned prow = A.n2 + 2 * kPad;                   // rows per slice
    const size_t plane = (size_t)prow * pitch * 2;           // floats per slice
    const float hi1 = (float)A.n1 + 2.5f, hi2 = (float)A.n2 + 2.5f;

    // boxes of chunk [cs, cs + kFpSlices) into box[slot] (threads < 64)
    auto make_box = [&](int cs, int slot) {
        if (tid >= 64) return;
        const int k = tid & (kFpSlices - 1);
        const int sl = cs + k;
        float mn1 = INFINITY, mx1 = -INFINITY, mn2 = INFINITY, mx2 = -INFINITY;
        if (sl <= ghi) {
            const float sf = (float)sl;
            for (int j = 0; j < kFpAngles; ++j)
                for (int q = 0; q < 4; ++q) {
                    if (!crng[j][q]) continue;
                    const float p1 = clamp_pad(fmaf(cpar[j][q][0], sf, cpar[j][q][1]), hi1);
                    const float p2 = clamp_pad(fmaf(cpar[j][q][2], sf, cpar[j][q][3]), hi2);
                    mn1 = fminf(mn1, p1);
                    mx1 = fmaxf(mx1, p1);
                    mn2 = fminf(mn2, p2);
                    mx2 = fmaxf(mx2, p2);
                }
        }
        int u0 = 0, v0 = 0, w = 0, h = 0;
        if (mn1 <= mx1) {
            u0 = max(0, (int)mn1 - 1);
            const int u1 = min((int)pitch - 2, (int)mx1 + 1);
            v0 = max(0, (int)mn2 - 1);
            const int v1 = min((int)prow - 1, (int)mx2 + 1);
            w = u1 - u0 + 1;
            h = v1 - v0 + 1;
        }
        // common width / height over the chunk's slices (lanes k < kFpSlices)
        int W = w, H = h;
        for (int o = 1; o < kFpSlices; o <<= 1) {
            W = max(W, __shfl_xor(W, o, 64));
            H = max(H, __shfl_xor(H, o, 64));
        }
        if (tid < kFpSlices) {
            box[slot].u0[k] = u0;
            box[slot].v0[k] = v0;
            if (tid == 0) {
                const int lp = W | 1;
                const bool fits = W > 0 && lp * H * kFpSlices <= kFpQuadCap;
                box[slot].w = fits ? W : 0;
                box[slot].h = H;
                box[slot].lp = lp;
            }
        }
    };

    float4 pre[kFpSlots];
    auto fetch = [&](int cs, const FpBox &b) {
        const int per_slice = b.w * b.h;
        const int n = per_slice * kFpSlices;
        const float inv_ps = 1.f / (float)per_slice, inv_w = 1.f / (float)b.w;
#pragma unroll
        for (int m = 0; m < kFpSlots; ++m) {
            const int i = (int)tid + 256 * m;
            if (i < n) {
                // exact for i < 2^12 (see the BP footprint loads)
                const int k = (int)(((float)i + 0.5f) * inv_ps);
                const int e = i - k * per_slice;
                const int r = (int)(((float)e + 0.5f) * inv_w);
                const int c = e - r * b.w;
                // a slice whose own box is smaller than the common one
                // reads clamped (unused) quads past its box
                const int sl = min(cs + k, (int)A.n0 - 1);
                const int rr = min(b.v0[k] + r, (int)prow - 1);
                const int cc = min(b.u0[k] + c, (int)pitch - 2);
                const float *pl = A.volp + (size_t)sl * plane;
                pre[m] = ld4b(pl, (unsigned)(rr * (int)pitch + cc) * 8u);
            }
        }
    };

    float acc0[kFpAngles], acc1[kFpAngles];
#pragma unroll
    for (int j = 0; j < kFpAngles; ++j) acc0[j] = acc1[j] = 0.f;

    make_box(glo, 0);
    __syncthreads();
    if (box[0].w) fetch(glo, box[0]);
    int slot = 0;
    for (int cs = glo; cs <= ghi; cs += kFpSlices, slot ^= 1) {
        const FpBox &b = box[slot];
        make_box(cs + kFpSlices, slot ^ 1);
        if (b.w) {
            const int per_slice = b.w * b.h;
            const int n = per_slice * kFpSlices;
            const float inv_ps = 1.f / (float)per_slice, inv_w = 1.f / (float)b.w;
#pragma unroll
            for (int m = 0; m < kFpSlots; ++m) {
                const int i = (int)tid + 256 * m;
                if (i < n) {
                    const int k = (int)(((float)i + 0.5f) * inv_ps);
                    const int e = i - k * per_slice;
                    const int r = (int)(((float)e + 0.5f) * inv_w);
                    tile[(k * b.h + r) * b.lp + (e - r * b.w)] = pre[m];
                }
            }
        }
        __syncthreads();  // tile and box[slot ^ 1] complete
        const FpBox &nb = box[slot ^ 1];
        if (cs + kFpSlices <= ghi && nb.w) fetch(cs + kFpSlices, nb);
        const int ce = min(cs + kFpSlices - 1, ghi);
        for (int sl = cs; sl <= ce; ++sl) {
            const int k = sl - cs;
            const float sf = (float)sl;
            const float *pl = A.volp + (size_t)sl * plane;
            const int rb = k * b.h, u0 = b.u0[k], v0 = b.v0[k];
#pragma unroll
            for (int j = 0; j < kFpAngles; ++j) {
                if (sl < slo[j] || sl > shi[j]) continue;  // wave-uniform
                const float p1 = clamp_pad(fmaf(a1[j], sf, c1[j]), hi1);
                const float p2 = clamp_pad(fmaf(a2[j], sf, c2[j]), hi2);
                float4 q;
                if (b.w) {
                    int off = ((int)(unsigned)p2 - v0 + rb) * b.lp + (int)(unsigned)p1 - u0;
                    // in range for every ray of the group by the box's
                    // margin; the clamp keeps the idle lanes (pixels past
                    // the detector) inside the tile
                    off = min(max(off, 0), kFpQuadCap - 1);
                    q = tile[off];
                } else {
                    q = ld4b(pl, pair_off(p1, p2, pitch));
                }
                const float x = bilerp_pairs(q, frac(p1), frac(p2));
                if ((sl - slo[j]) & 1)
                    acc1[j] += x;
                else
                    acc0[j] += x;
            }
        }
        __syncthreads();  // the tile is read; box[slot] may be rewritten
    }
    if (!pix) return;
#pragma unroll
    for (int j = 0; j < kFpAngles; ++j) {
        if (angle[j] < 0) continue;
        const float val = (acc0[j] + acc1[j]) * sc[j];
        const int an = angle[j];
        size_t idx = ((size_t)v * A.na + an) * A.nu + u;
        size_t oidx = A.padded_out
                          ? ((size_t)an * (A.nv + 2 * kPad) + (v + kPad)) * (A.nu + 2 * kPad) + u + kPad
                          : idx;
        if (A.mode == OUT_ADD) {
            A.out[oidx] += val;
        } else if (A.mode == OUT_SET) {
            A.out[oidx] = val;
        } else if (A.mode == OUT_RESID) {
            float r = A.rb[idx] + val;
            if (A.rw) r *= A.rw[oidx];
            A.out[oidx] = r;
        } else if (A.mode == OUT_RATIO) {
            A.out[oidx] = em_ratio(A.rb[idx], val);
        } else {
            float r = A.out[oidx] + val;
            if (A.rw) r *= A.rw[oidx];
            A.out[oidx] = r;
        }
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
    fp3d_ray<GEO_PAR3D>(A, A.ang[angle], u, v, a1f, a2f, c1f, c2f, scale, lo, hi);
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
    const float *pre; // bp3d_kernel: BP of earlier angle chunks, added before the epilogue (or NULL)
};

// Epilogue of a column of ZS voxels: scale, add the earlier chunks, write
// in the launch's mode.
// Every load of the column (earlier chunks, old values, weights) is issued
// before the first store, so the column's read-modify-write costs one
// memory round trip, not one per voxel (the compiler cannot reorder a load
// of vol after a store to vol by itself).  Z0 is the workgroup's, so the
// bounds tests are uniform.
template <int ZS>
__device__ __forceinline__ void bp_column_store(const BPArgs &A, int X, int Y, int Z0,
                                                const float (&acc)[ZS]) {
    const int nzr = min(ZS, (int)A.nz - Z0);
    const size_t plane = (size_t)A.ny * A.nx;
    const size_t base = ((size_t)Z0 * A.ny + Y) * A.nx + X;
    const bool old = A.mode != OUT_SET;
    const bool wts = (A.mode == OUT_EMUPD || A.mode == OUT_UPDATE) && A.uw;
    float pv[ZS], ov[ZS], wv[ZS];
#pragma unroll
    for (int i = 0; i < ZS; ++i) {
        pv[i] = 0.f;
        ov[i] = 0.f;
        wv[i] = 1.f;
        if (i < nzr) {
            const size_t idx = base + i * plane;
            if (A.pre) pv[i] = A.pre[idx];
            if (old) ov[i] = A.vol[idx];
            if (wts) wv[i] = A.uw[idx];
        }
    }
#pragma unroll
    for (int i = 0; i < ZS; ++i) {
        if (i < nzr) {
            float val = acc[i] * A.osc;
            if (A.pre) val += pv[i];
            float out;
            if (A.mode == OUT_ADD) out = ov[i] + val;
            else if (A.mode == OUT_SET) out = val;
            else if (A.mode == OUT_EMUPD) out = ov[i] * (val * wv[i]);
            else out = fminf(fmaxf(fmaf(wv[i], val, ov[i]), A.lo), A.hi);  // OUT_UPDATE
            A.vol[base + i * plane] = out;
        }
    }
}

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
            const float wu = frac(pu);
            const unsigned ou = (unsigned)pu << 3;
            const float wgt = (KIND == 0) ? fr * fr : (KIND == 1 ? fr : wa);
            // every tap load of the column issued before the first use: the
            // loads' latency, not their issue rate, bounds this loop
            float4 q[ZS];
            float wv[ZS];
#pragma unroll
            for (int i = 0; i < ZS; ++i) {
                // pv in [0.5, nv + 2.5]: truncation is the floor
                const float pv = clamp_pad(vn * fr, hiv);
                wv[i] = frac(pv);
                q[i] = ld4b(pl, __umul24((unsigned)pv, rowb) + ou);
                vn += c.nv[2];
            }
            // keep the scheduler from sinking the loads into their uses
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int i = 0; i < ZS; ++i) acc[i] = fmaf(wgt, bilerp_pairs(q[i], wu, wv[i]), acc[i]);
        } else {
#pragma unroll
            for (int i = 0; i < ZS; ++i) {
                float fr = (KIND == 2) ? 1.f : __builtin_amdgcn_rcpf(dn);
                float pu = clamp_pad(un * fr, hiu);
                float pv = clamp_pad(vn * fr, hiv);
                float val = bilerp_pairs(ld4b(pl, pair_off(pu, pv, pitch)), frac(pu), frac(pv));
                float wgt = (KIND == 0) ? fr * fr : (KIND == 1 ? fr : wa);
                acc[i] = fmaf(wgt, val, acc[i]);
                un += c.nu[2];
                vn += c.nv[2];
                dn += c.dn[2];
            }
        }
    }
    bp_column_store<ZS>(A, X, Y, Z0, acc);
}

// ---- BP with the detector footprint staged in LDS -----------------------
// The z-constant BP above reads one 16-byte tap (two difference pairs) per
// voxel and angle through the texture-address path, which bounds it.  Here
// each workgroup (16 x 16 columns of ZS voxels) stages, per angle, the
// rectangle of the pair plane that its block projects into as 16-byte
// quads -- quad (r, c) = pairs (r, c) and (r, c + 1), the whole tap -- so
// that every tap is one aligned 16-byte LDS read.  Thread t stages column
// t % 64 of rows t / 64 + 4 k (k < 8) with 16-byte loads issued one angle
// ahead and held in registers while the current angle is sampled, so the
// staging needs no index division.  The footprint is the bounding box of
// the projected corners of the block (u and v are linear-fractional in the
// voxel position, so the extremes over the block are at its corners; the
// clamp to the padded range is monotone), widened by one texel on each
// side, far beyond the fp32 rounding of the taps (replayed on the CPU over
// whole geometries by the tests).  Taps, weights and the order of the sums
// are those of bp3d_kernel<KIND, ZS, true>, so both give the same floats.
// An angle whose footprint exceeds 64 quads x 32 rows or the tile, or whose
// denominator changes sign over the block, is sampled from global memory
// as in bp3d_kernel.  A staged row is `lp` quads long: lp = w | 1 (pm < 0)
// or w + ((pm - w) mod 16) (pm in [0, 16)), so that rows start at
// different offsets modulo the 16 quads (256 bytes) of one LDS cycle.
constexpr int kTileCap = 2048;   // LDS quads of the staged tile (32 KiB)
constexpr int kFpGroup = 256;    // angles whose footprints are held at once
constexpr int kStageCols = 64;   // widest staged footprint (quads)
constexpr int kStageRows = 32;   // tallest staged footprint (rows): 4 per thread slot
constexpr int kStageSlots = kStageRows / 4;

__device__ __forceinline__ int lds_row_pitch(int w, int pm) {
    return pm < 0 ? (w | 1) : w + ((pm - w) & 15);
}

template <int KIND>
__device__ int4 bp_footprint(const BP3DAngle &c, float x0, float x1, float y0, float y1, float z0,
                             float z1, float hiu, float hiv, int pitch, int prow, int pm) {
    float umin = INFINITY, umax = -INFINITY, vmin = INFINITY, vmax = -INFINITY;
    float dmin = INFINITY, dmax = -INFINITY;
    for (int j = 0; j < 4; ++j) {
        const float x = (j & 1) ? x1 : x0, y = (j & 2) ? y1 : y0;
        const float un = fmaf(y, c.nu[1], fmaf(x, c.nu[0], c.nu[3]));
        const float dn = fmaf(y, c.dn[1], fmaf(x, c.dn[0], c.dn[3]));
        dmin = fminf(dmin, dn);
        dmax = fmaxf(dmax, dn);
        const float fr = (KIND == 2) ? 1.f : 1.f / dn;
        const float pu = clamp_pad(un * fr, hiu);
        umin = fminf(umin, pu);
        umax = fmaxf(umax, pu);
        for (int k = 0; k < 2; ++k) {
            const float z = k ? z1 : z0;
            const float vn = fmaf(z, c.nv[2], fmaf(y, c.nv[1], fmaf(x, c.nv[0], c.nv[3])));
            const float pv = clamp_pad(vn * fr, hiv);
            vmin = fminf(vmin, pv);
            vmax = fmaxf(vmax, pv);
        }
    }
    const bool sign_ok = (KIND == 2) || dmin > 0.f || dmax < 0.f;
    const bool finite = __builtin_isfinite(umin) && __builtin_isfinite(umax) &&
                        __builtin_isfinite(vmin) && __builtin_isfinite(vmax);
    if (!sign_ok || !finite) return make_int4(0, 0, 0, 0);
    // taps: quad floor(pu) (pairs floor(pu), floor(pu) + 1) of pair row
    // floor(pv); one texel of margin on each side.  Quad c reads pairs c and
    // c + 1, so the last quad column is at most pitch - 2.
    const int u0 = max(0, (int)umin - 1), u1 = min(pitch - 2, (int)umax + 1);
    const int v0 = max(0, (int)vmin - 1), v1 = min(prow - 1, (int)vmax + 1);
    const int w = u1 - u0 + 1, h = v1 - v0 + 1;
    if (w < 1 || w > kStageCols || h > kStageRows || lds_row_pitch(w, pm) * h > kTileCap)
        return make_int4(0, 0, 0, 0);
    return make_int4(u0, v0, w, h);
}

template <int KIND, int ZS>
__global__ __launch_bounds__(256) void bp3d_lds_kernel(BPArgs A, int pm) {
    __shared__ float4 tile[kTileCap];
    __shared__ int4 fps[kFpGroup];
    const unsigned tid = threadIdx.x;
    const unsigned tiles_x = (A.nx + 15) / 16;
    const int X0 = (blockIdx.x % tiles_x) * 16, Y0 = (blockIdx.x / tiles_x) * 16;
    const int X = X0 + (int)(tid & 15);
    const int Y = Y0 + (int)(tid >> 4);
    const int Z0 = blockIdx.y * ZS;
    const bool active = X < (int)A.nx && Y < (int)A.ny;
    const float hx = 0.5f * A.nx - 0.5f, hy = 0.5f * A.ny - 0.5f, hz = 0.5f * A.nz - 0.5f;
    const float fx = X - 0.5f * A.nx + 0.5f;
    const float fy = Y - 0.5f * A.ny + 0.5f;
    const float fz = Z0 - 0.5f * A.nz + 0.5f;
    // corners of the block's voxel centres (the columns inside the volume,
    // all ZS slices: the slices past nz are computed and discarded)
    const float cx0 = X0 - hx, cx1 = min(X0 + 15, (int)A.nx - 1) - hx;
    const float cy0 = Y0 - hy, cy1 = min(Y0 + 15, (int)A.ny - 1) - hy;
    const float cz0 = Z0 - hz, cz1 = (Z0 + ZS - 1) - hz;
    const unsigned pitch = A.nu + 2 * kPad;                    // pairs per row
    const unsigned prow = A.nv + 2 * kPad;                     // rows per angle
    const size_t plane = (size_t)prow * pitch;                 // pairs per angle
    const unsigned rowb = pitch * 8u;
    const float hiu = (float)A.nu + 2.5f, hiv = (float)A.nv + 2.5f;
    const int sc = (int)(tid & 63), sr = (int)(tid >> 6);       // staging column, first row
    const char *tile_b = reinterpret_cast<const char *>(tile);
    float acc[ZS];
#pragma unroll
    for (int i = 0; i < ZS; ++i) acc[i] = 0.f;
    float4 pre[kStageSlots];

    // this thread's quads of the footprint: column sc of rows sr + 4 k
    auto fetch = [&](unsigned a, int4 f) {
        const float *pl = A.pp + (size_t)a * plane * 2;
        unsigned off = ((unsigned)(f.y + sr) * pitch + (unsigned)(f.x + sc)) * 8u;
        const unsigned step = 4u * rowb;
#pragma unroll
        for (int k = 0; k < kStageSlots; ++k, off += step)
            if (sc < f.z && sr + 4 * k < f.w) pre[k] = ld4b(pl, off);
    };

    for (unsigned g0 = 0; g0 < A.na; g0 += kFpGroup) {
        const unsigned gn = min((unsigned)kFpGroup, A.na - g0);
        __syncthreads();  // the previous group's footprints and tile are no longer read
        if (tid < gn)
            fps[tid] = bp_footprint<KIND>(A.ang[g0 + tid], cx0, cx1, cy0, cy1, cz0, cz1, hiu, hiv,
                                          (int)pitch, (int)prow, pm);
        __syncthreads();
        if (fps[0].z) fetch(g0, fps[0]);
        for (unsigned k = 0; k < gn; ++k) {
            const unsigned a = g0 + k;
            const int4 f = fps[k];
            const int lp = lds_row_pitch(f.z, pm);
            if (k) __syncthreads();  // every thread is done with angle a - 1's tile
            if (f.z) {
                int idx = sr * lp + sc;
#pragma unroll
                for (int j = 0; j < kStageSlots; ++j, idx += 4 * lp)
                    if (sc < f.z && sr + 4 * j < f.w) tile[idx] = pre[j];
            }
            __syncthreads();  // the tile holds angle a
            if (k + 1 < gn) {
                const int4 fn = fps[k + 1];
                if (fn.z) fetch(a + 1, fn);
            }
            if (!active) continue;
            const BP3DAngle c = A.ang[a];
            const float un = fmaf(fz, c.nu[2], fmaf(fy, c.nu[1], fmaf(fx, c.nu[0], c.nu[3])));
            float vn = fmaf(fz, c.nv[2], fmaf(fy, c.nv[1], fmaf(fx, c.nv[0], c.nv[3])));
            const float dn = fmaf(fz, c.dn[2], fmaf(fy, c.dn[1], fmaf(fx, c.dn[0], c.dn[3])));
            const float wa = (KIND == 2) ? A.aw[a] : 1.f;
            const float fr = (KIND == 2) ? 1.f : __builtin_amdgcn_rcpf(dn);
            const float pu = clamp_pad(un * fr, hiu);
            const float wu = frac(pu);
            const float wgt = (KIND == 0) ? fr * fr : (KIND == 1 ? fr : wa);
            if (f.z) {
                // byte offset of tap row r: r * lp * 16 + (column - u0 - v0 lp) * 16;
                // inside the tile by the footprint's margin
                const int lp16 = lp * 16;
                const int base = ((int)(unsigned)pu - f.x - f.y * lp) * 16;
                float4 q[ZS];
                float wv[ZS];
#pragma unroll
                for (int i = 0; i < ZS; ++i) {
                    const float pv = clamp_pad(vn * fr, hiv);
                    wv[i] = frac(pv);
                    const int off = __mul24((int)(unsigned)pv, lp16) + base;
                    q[i] = *reinterpret_cast<const float4 *>(tile_b + off);
                    vn += c.nv[2];
                }
#pragma unroll
                for (int i = 0; i < ZS; ++i) acc[i] = fmaf(wgt, bilerp_pairs(q[i], wu, wv[i]), acc[i]);
            } else {
                const float *pl = A.pp + (size_t)a * plane * 2;
                const unsigned ou = (unsigned)pu << 3;
                float4 q[ZS];
                float wv[ZS];
#pragma unroll
                for (int i = 0; i < ZS; ++i) {
                    const float pv = clamp_pad(vn * fr, hiv);
                    wv[i] = frac(pv);
                    q[i] = ld4b(pl, __umul24((unsigned)pv, rowb) + ou);
                    vn += c.nv[2];
                }
#pragma unroll
                for (int i = 0; i < ZS; ++i) acc[i] = fmaf(wgt, bilerp_pairs(q[i], wu, wv[i]), acc[i]);
            }
        }
    }
    if (active) bp_column_store<ZS>(A, X, Y, Z0, acc);
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
                    float val = bilerp_pairs(ld4b(pl, pair_off(pu, pv, pitch)), frac(pu), frac(pv));
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

// ------------------------------------------------ cylindrical BP --------
// Voxel-driven BP from a cylindrical detector (dev_cylcone_BP,
// cuda/3d/cone_cyl_bp.cu:58-158), in world units of the translated frame:
// the ray r from the source through the voxel, in the (radial, tangential,
// axial) frame, meets the cylinder at the larger root t of
// A t^2 + 2 B t + C = 0 (A = ra^2 + rb^2, B = ra sa + rb sb); the detector
// coordinates there are u = asin((t rb + sb) / R) / angular step + nu / 2
// and v = (t rc + sc) / |v| + nv / 2 (texture coordinates); weight t^2.
// A voxel whose ray misses the cylinder (B^2 < A C) gets nothing from that
// angle (its coordinates are replaced by an in-bounds dummy).  One thread
// per (x, y) column of ZS voxels, 16 x 16 threads per group, as bp3d_kernel.
template <int ZS>
__global__ __launch_bounds__(256) void bp3d_cyl_kernel(BPArgs A, const CylBPAngle *__restrict__ ang,
                                                       float vx, float vy, float vz) {
    const unsigned tiles_x = (A.nx + 15) / 16;
    const int X = (blockIdx.x % tiles_x) * 16 + (threadIdx.x & 15);
    const int Y = (blockIdx.x / tiles_x) * 16 + (threadIdx.x >> 4);
    const int Z0 = blockIdx.y * ZS;
    if (X >= (int)A.nx || Y >= (int)A.ny) return;
    const float fx = vx * (X - 0.5f * A.nx + 0.5f);
    const float fy = vy * (Y - 0.5f * A.ny + 0.5f);
    const float fz = vz * (Z0 - 0.5f * A.nz + 0.5f);
    const unsigned pitch = A.nu + 2 * kPad;
    const size_t plane = (size_t)(A.nv + 2 * kPad) * pitch * 2;
    const float hiu = (float)A.nu + 2.5f, hiv = (float)A.nv + 2.5f;
    // texture coordinate -> padded texel coordinate: - 0.5 + 2
    const float ou = 0.5f * A.nu + 1.5f, ov = 0.5f * A.nv + 1.5f;
    float acc[ZS];
#pragma unroll
    for (int i = 0; i < ZS; ++i) acc[i] = 0.f;
    const float *pl = A.pp;
    for (unsigned a = 0; a < A.na; ++a, pl += plane) {
        const CylBPAngle c = ang[a];
        float ra = fmaf(fz, c.ca[2], fmaf(fy, c.ca[1], fmaf(fx, c.ca[0], c.ca[3])));
        float rb = fmaf(fz, c.cb[2], fmaf(fy, c.cb[1], fmaf(fx, c.cb[0], c.cb[3])));
        float rc = fmaf(fz, c.cc[2], fmaf(fy, c.cc[1], fmaf(fx, c.cc[0], c.cc[3])));
        const float dra = vz * c.ca[2], drb = vz * c.cb[2], drc = vz * c.cc[2];
#pragma unroll
        for (int i = 0; i < ZS; ++i) {
            const float qa = fmaf(ra, ra, rb * rb);
            const float qb = fmaf(ra, c.sa, rb * c.sb);
            const float disc = fmaf(qb, qb, -qa * c.c);
            const bool hit = disc >= 0.f;
            const float t = hit ? (sqrtf(disc) - qb) / qa : 0.f;
            const float sn = fminf(fmaxf(fmaf(t, rb, c.sb) * c.inv_r, -1.f), 1.f);
            const float pu = hit ? clamp_pad(fmaf(asinf(sn), c.inv_dth, ou), hiu) : 0.5f;
            const float pv = hit ? clamp_pad(fmaf(fmaf(t, rc, c.sc), c.inv_nv, ov), hiv) : 0.5f;
            const float val = bilerp_pairs(ld4b(pl, pair_off(pu, pv, pitch)), frac(pu), frac(pv));
            acc[i] = fmaf(t * t, val, acc[i]);
            ra += dra;
            rb += drb;
            rc += drc;
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

// --------------------------------------------------- matched BP ----------
// Exact transpose of fp3d_kernel, voxel-driven (a gather, no atomics).  The
// forward projector takes, per (sub-)ray r and slice s, the bilinear
// sample at padded in-slice coordinates p(r, s) = (a1 s + c1, a2 s + c2)
// (fp32).  Its transpose at the voxel with padded in-slice indices
// (J1, J2) of slice s is
//     sum_r y_r * scale_r / R^2 * (1 - |p1 - J1|)+ * (1 - |p2 - J2|)+ ,
// over the rays of the voxel's direction group whose sample lies within
// one texel of the voxel along both axes.  Interior voxels only receive
// samples with unclamped coordinates inside the ray's clipped slice range,
// so clamping and clipping need no counterpart here.
//
// matched_rays_kernel stores, per (sub-)ray of a chunk of angles, the fp32
// coefficients (a1, c1, a2, c2) exactly as fp3d_ray computes them for the
// forward projector, and the weighted value y * scale / R^2.  Sub-rays
// form a lattice (U, V) = (u R + k, v R + l).
template <int GK>
__global__ __launch_bounds__(256) void matched_rays_kernel(FPArgs A, const float *__restrict__ proj,
                                                           float4 *__restrict__ rays,
                                                           float *__restrict__ yw) {
    const unsigned R = A.ss, NU = A.nu * R, NV = A.nv * R;
    const size_t per = (size_t)NU * NV;
    const size_t n = per * A.lcount;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const unsigned li = (unsigned)(i / per);
        const unsigned rem = (unsigned)(i - (size_t)li * per);
        const unsigned V = rem / NU, U = rem - V * NU;
        const int angle = A.alist[A.lfirst + li];
        const int u = (int)(U / R), k = (int)(U % R), v = (int)(V / R), l = (int)(V % R);
        const FP3DAngle &G = A.ang[angle];
        float a1f, a2f, c1f, c2f, scale;
        int lo, hi;
        if (R == 1)
            fp3d_ray<GK>(A, G, u, v, a1f, a2f, c1f, c2f, scale, lo, hi);
        else
            fp3d_ray<GK>(A, G, u - 0.5 + (k + 0.5) / (int)R, v - 0.5 + (l + 0.5) / (int)R, a1f,
                           a2f, c1f, c2f, scale, lo, hi);
        rays[i] = make_float4(a1f, c1f, a2f, c2f);
        const float y = proj[((size_t)v * A.na + angle) * A.nu + u];
        yw[i] = y * scale / (float)(R * R);
    }
}

__device__ __forceinline__ void cross3(const double *a, const double *b, double *c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ __forceinline__ double dot3(const double *a, const double *b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// Per angle of the chunk, the projective map from a point W of the group
// frame (c0, c1, c2; voxel units, volume centre at 0) to continuous
// sub-ray lattice coordinates: U = nU.(W, 1) / den.(W, 1), likewise V;
// three float4 (nU, nV, den), scaled so that den is 1 at the centre.
// Cone: W = S + t (P - S) with P = p0 + u U + v V on the detector gives,
// with n = U x V and k = (p0 - S).n,
//     u = (W - S).Au / (|n|^2 (W - S).n),  Au = ((S - p0).(V x n)) n + k V x n,
// and v likewise with n x U.  Parallel: W - p0 = u U + v V + t d gives
//     u = (W - p0).(V x d) / U.(V x d),  v = (W - p0).(d x U) / V.(d x U).
template <bool CONE>
__global__ void matched_geo_kernel(FPArgs A, float4 *__restrict__ geo) {
    const unsigned li = blockIdx.x * blockDim.x + threadIdx.x;
    if (li >= A.lcount) return;
    const FP3DAngle &G = A.ang[A.alist[A.lfirst + li]];
    double nu4[4], nv4[4], dn4[4];
    if (CONE) {
        double n[3], mu[3], mv[3], sp[3], ps[3], Au[3], Av[3];
        cross3(G.u, G.v, n);
        cross3(G.v, n, mu);
        cross3(n, G.u, mv);
        for (int i = 0; i < 3; ++i) {
            sp[i] = G.s[i] - G.p0[i];
            ps[i] = -sp[i];
        }
        const double k = dot3(ps, n), nn = dot3(n, n);
        const double su = dot3(sp, mu), sv = dot3(sp, mv);
        for (int i = 0; i < 3; ++i) {
            Au[i] = su * n[i] + k * mu[i];
            Av[i] = sv * n[i] + k * mv[i];
        }
        for (int i = 0; i < 3; ++i) {
            nu4[i] = Au[i];
            nv4[i] = Av[i];
            dn4[i] = nn * n[i];
        }
        nu4[3] = -dot3(G.s, Au);
        nv4[3] = -dot3(G.s, Av);
        dn4[3] = -nn * dot3(G.s, n);
    } else {
        double bu[3], bv[3];
        cross3(G.v, G.s, bu);
        cross3(G.s, G.u, bv);
        const double iu = 1.0 / dot3(G.u, bu), iv = 1.0 / dot3(G.v, bv);
        for (int i = 0; i < 3; ++i) {
            nu4[i] = bu[i] * iu;
            nv4[i] = bv[i] * iv;
            dn4[i] = 0.0;
        }
        nu4[3] = -dot3(G.p0, bu) * iu;
        nv4[3] = -dot3(G.p0, bv) * iv;
        dn4[3] = 1.0;
    }
    // detector coordinate u -> lattice coordinate R u + (R - 1) / 2
    const double R = (double)A.ss, off = 0.5 * (R - 1.0);
    const double inv = 1.0 / dn4[3];
    float4 *o = geo + 3 * (size_t)li;
    double r[3][4];
    for (int i = 0; i < 4; ++i) {
        r[0][i] = (R * nu4[i] + off * dn4[i]) * inv;
        r[1][i] = (R * nv4[i] + off * dn4[i]) * inv;
        r[2][i] = dn4[i] * inv;
    }
    for (int j = 0; j < 3; ++j)
        o[j] = make_float4((float)r[j][0], (float)r[j][1], (float)r[j][2], (float)r[j][3]);
}

struct MatchedArgs {
    const float4 *rays;  // [angle of chunk][V][U] (a1, c1, a2, c2)
    const float *yw;     // [angle of chunk][V][U] y * scale / R^2
    const float4 *geo;   // [angle of chunk][3] lattice map (flat detector)
    const CylBPAngle *cyl;  // cylindrical detector: per-angle map (world units)
    const int *alist;       // cylindrical detector: angle indices of the chunk
    float vs[3];            // cylindrical detector: voxel sizes (x, y, z)
    unsigned count;      // angles in the chunk
    unsigned NU, NV;     // sub-ray lattice size
    unsigned n0, n1, n2, tiles1;
    unsigned nx, ny;
    int group;
    float *vol;  // [z][y][x], += the chunk's contribution
};

// One thread per S consecutive voxels along the slice axis c0 (S slices
// share one pass over their candidate rays: every candidate is loaded once
// and weighted for each slice).  A workgroup is a 16 (c1) x 16 (c2) tile
// of S slices; consecutive workgroups take consecutive slice groups of the
// same tile (their candidate rays overlap).  Per angle the candidate rays
// are the lattice points inside the image of the voxels' open
// neighbourhood squares (c1 +- 1, c2 +- 1) over the S slices, bounded by
// the projected corners of the first and last slice's squares (a central
// or parallel projection maps each corner's path to a segment), widened by
// 0.05 lattice units for fp32 rounding of the map; extra candidates get
// weight 0.
// CYL: cylindrical detector; the corners map through the cylinder hit of
// bp3d_cyl_kernel (world units), which bends the square's edges slightly,
// so the bound is widened by 0.25 lattice units, and a corner whose ray
// misses the cylinder makes the scan cover the whole detector.
template <int S, bool CYL>
__global__ __launch_bounds__(256) void bp3d_matched_gather(MatchedArgs M) {
    const unsigned sgroups = (M.n0 + S - 1) / S;
    const unsigned s0 = (blockIdx.x % sgroups) * S, tile = blockIdx.x / sgroups;
    const int i1 = (int)((tile % M.tiles1) * 16 + (threadIdx.x & 15));
    const int i2 = (int)((tile / M.tiles1) * 16 + (threadIdx.x >> 4));
    if (i1 >= (int)M.n1 || i2 >= (int)M.n2) return;
    const float w0 = (float)s0 - 0.5f * M.n0 + 0.5f;
    const float w1 = (float)i1 - 0.5f * M.n1 + 0.5f;
    const float w2 = (float)i2 - 0.5f * M.n2 + 0.5f;
    const float J1 = (float)(i1 + kPad), J2 = (float)(i2 + kPad), sf = (float)s0;
    const float NUm = (float)M.NU - 1.f, NVm = (float)M.NV - 1.f;
    constexpr float kEps = CYL ? 0.25f : 0.05f;
    float acc[S];
#pragma unroll
    for (int k = 0; k < S; ++k) acc[k] = 0.f;
    for (unsigned a = 0; a < M.count; ++a) {
        float umin = INFINITY, umax = -INFINITY, vmin = INFINITY, vmax = -INFINITY;
        if (CYL) {
            const CylBPAngle c = M.cyl[M.alist[a]];
            bool miss = false;
#pragma unroll
            for (int q = 0; q < (S > 1 ? 8 : 4); ++q) {
                const float g0 = w0 + ((q & 4) ? (float)(S - 1) : 0.f);
                const float g1 = w1 + ((q & 1) ? 1.f : -1.f), g2 = w2 + ((q & 2) ? 1.f : -1.f);
                float x, y, z;
                if (M.group == 0) { x = g0; y = g1; z = g2; }
                else if (M.group == 1) { y = g0; x = g1; z = g2; }
                else { z = g0; x = g1; y = g2; }
                x *= M.vs[0];
                y *= M.vs[1];
                z *= M.vs[2];
                const float ra = fmaf(z, c.ca[2], fmaf(y, c.ca[1], fmaf(x, c.ca[0], c.ca[3])));
                const float rb = fmaf(z, c.cb[2], fmaf(y, c.cb[1], fmaf(x, c.cb[0], c.cb[3])));
                const float rc = fmaf(z, c.cc[2], fmaf(y, c.cc[1], fmaf(x, c.cc[0], c.cc[3])));
                const float qa = fmaf(ra, ra, rb * rb);
                const float qb = fmaf(ra, c.sa, rb * c.sb);
                const float disc = fmaf(qb, qb, -qa * c.c);
                miss |= !(disc >= 0.f);
                const float t = (sqrtf(fmaxf(disc, 0.f)) - qb) / qa;
                const float sn = fminf(fmaxf(fmaf(t, rb, c.sb) * c.inv_r, -1.f), 1.f);
                // lattice coordinate = texture coordinate - 0.5
                const float U = fmaf(asinf(sn), c.inv_dth, 0.5f * (M.NU - 1.f));
                const float V = fmaf(fmaf(t, rc, c.sc), c.inv_nv, 0.5f * (M.NV - 1.f));
                umin = fminf(umin, U);
                umax = fmaxf(umax, U);
                vmin = fminf(vmin, V);
                vmax = fmaxf(vmax, V);
            }
            if (miss) {
                umin = vmin = 0.f;
                umax = NUm;
                vmax = NVm;
            }
        } else {
            const float4 gu = M.geo[3 * a], gv = M.geo[3 * a + 1], gd = M.geo[3 * a + 2];
            const float cu = fmaf(gu.x, w0, fmaf(gu.y, w1, fmaf(gu.z, w2, gu.w)));
            const float cv = fmaf(gv.x, w0, fmaf(gv.y, w1, fmaf(gv.z, w2, gv.w)));
            const float cd = fmaf(gd.x, w0, fmaf(gd.y, w1, fmaf(gd.z, w2, gd.w)));
#pragma unroll
            for (int c = 0; c < (S > 1 ? 8 : 4); ++c) {
                const float d0 = (c & 4) ? (float)(S - 1) : 0.f;
                const float d1 = (c & 1) ? 1.f : -1.f, d2 = (c & 2) ? 1.f : -1.f;
                const float r = __builtin_amdgcn_rcpf(
                    fmaf(d2, gd.z, fmaf(d1, gd.y, fmaf(d0, gd.x, cd))));
                const float U = fmaf(d2, gu.z, fmaf(d1, gu.y, fmaf(d0, gu.x, cu))) * r;
                const float V = fmaf(d2, gv.z, fmaf(d1, gv.y, fmaf(d0, gv.x, cv))) * r;
                umin = fminf(umin, U);
                umax = fmaxf(umax, U);
                vmin = fminf(vmin, V);
                vmax = fmaxf(vmax, V);
            }
        }
        // clamped in float before the conversion: a NaN bound becomes a full
        // (in-range) scan, an infinite one an empty scan
        const int u0 = (int)fminf(fmaxf(ceilf(umin - kEps), 0.f), NUm + 1.f);
        const int u1 = (int)fmaxf(fminf(floorf(umax + kEps), NUm), -1.f);
        const int v0 = (int)fminf(fmaxf(ceilf(vmin - kEps), 0.f), NVm + 1.f);
        const int v1 = (int)fmaxf(fminf(floorf(vmax + kEps), NVm), -1.f);
        // per-angle tables at a scalar base; 32-bit byte offsets inside
        const size_t abase = (size_t)a * M.NV * M.NU;
        const float4 *ra = M.rays + abase;
        const float *ya = M.yw + abase;
        for (int V = v0; V <= v1; ++V) {
            const unsigned rb = (unsigned)V * M.NU;
            for (int U = u0; U <= u1; ++U) {
                const unsigned o = rb + (unsigned)U;
                const float4 r = ld4a(ra, o << 4);
                const float y = ld1b(ya, o << 2);
#pragma unroll
                for (int k = 0; k < S; ++k) {
                    const float sk = sf + (float)k;
                    const float e1 = tent(fmaf(r.x, sk, r.y) - J1);
                    const float e2 = tent(fmaf(r.z, sk, r.w) - J2);
                    acc[k] = fmaf(y * e1, e2, acc[k]);
                }
            }
        }
    }
#pragma unroll
    for (int k = 0; k < S; ++k) {
        const unsigned s = s0 + k;
        if (s >= M.n0) break;
        unsigned x, y, z;
        if (M.group == 0) { x = s; y = i1; z = i2; }
        else if (M.group == 1) { y = s; x = i1; z = i2; }
        else { z = s; x = i1; y = i2; }
        M.vol[((size_t)z * M.ny + y) * M.nx + x] += acc[k];
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
        dim3 grid((g.ny + 2 * kPad + 63) / 64, (g.nx + 63) / 64,
                  (g.nz + 2 * kPad + kTzRows - 1) / kTzRows);
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

// z voxels per thread of bp3d_kernel: 16, or 4 / 8 via ASTRA_AMD_BP3D_ZS.
// ASTRA_AMD_BP3D_LDS=1: the z-constant BP stages detector footprints in LDS
// (read at every launch)
bool bp_lds() {
    const char *e = std::getenv("ASTRA_AMD_BP3D_LDS");
    return e && e[0] == '1';
}

// LDS row pitch of the staged BP: ASTRA_AMD_BP3D_LDS_PITCH=<0..15> pads a
// staged row to that residue modulo 16 quads, -1 (default) to the next odd
// length (read at every launch)
int bp_lds_pitch_mod() {
    const char *e = std::getenv("ASTRA_AMD_BP3D_LDS_PITCH");
    const int x = e ? std::atoi(e) : -1;
    return x < 0 ? -1 : (x & 15);
}

// Measured at 512^3 / 720 x 768^2: 4: 52.7 ms, 8: 50.0, 16: 49.2 per BP
// launch; at 1024^3 / 1440 x 1536^2 (40-angle chunks) 16 is 1 % faster
// than 8.
int bp_zs() {
    static const int v = [] {
        const char *e = std::getenv("ASTRA_AMD_BP3D_ZS");
        const int x = e ? std::atoi(e) : 16;
        return (x == 4 || x == 8) ? x : 16;
    }();
    return v;
}

// Angles per bp3d_kernel launch.  A workgroup sweeps every angle of the
// launch over its voxels, and workgroups start whenever a slot frees, so in
// steady state the resident workgroups sit at angles spread evenly over the
// launch's range: with 1440 planes of 19 MB (1024^3 / 1536^2) two
// neighbouring tiles almost never read the same plane at the same time and
// the L2 hit rate of the BP falls to 8 %.  Launching the angles in chunks
// keeps the resident workgroups within fewer planes, at the price of one
// volume read-modify-write per chunk.  Measured at 1024^3 / 1440 x 1536^2,
// chunks of 20 / 40 / 80 / 160 / 1440 angles: 1034 / 1004 / 1035 / 1052 /
// 1070 ms per BP with 8 z voxels per thread (L2 hit rate 57 % at 40), 992
// / 1001 ms for 40 / 80 with 16; at 512^3 / 720 x 768^2 (3.4 GB of padded
// planes, 16 z voxels per thread), chunks of 90 / 180 / 240 / 360 / 720:
// 49.5 / 48.6 / 48.2 / 48.7 / 49.3 ms.  So: 40 angles from 8 GiB
// of padded projections on; below that, equal chunks of at most 1.2 GB of
// planes (a single launch below 1.2 GB); in every output mode.
// ASTRA_AMD_BP3D_ACHUNK=<angles> forces a chunk size (0: default rule; read
// at every call, so a process can change it between calls).
unsigned bp_achunk(size_t pair_bytes, unsigned na) {
    const char *e = std::getenv("ASTRA_AMD_BP3D_ACHUNK");
    const int forced = e ? std::atoi(e) : 0;
    if (forced > 0) return (unsigned)forced;
    if (pair_bytes >= (8ull << 30)) return 40u;
    const size_t cap = 1200000000ull;
    if (pair_bytes <= cap || na == 0) return std::max(na, 1u);
    const size_t chunks = (pair_bytes + cap - 1) / cap;
    return (unsigned)((na + chunks - 1) / chunks);
}

// Most angles per block of the FP workgroup map (fp_map): an XCD's range
// is cut into equal blocks of at most this many angles.
// ASTRA_AMD_FP3D_ANGLE_BLOCK overrides the default 64 (1..64; read at every
// call).  Measured at 512^3 / 720 x 768^2 (45 angles per XCD and group),
// ms per group launch: 16: 43.5, 23: 42.9, 32 (blocks of 32 and 13): 44.0,
// 45 or 64 (one block): 42.8, 42.7.
unsigned fp_angle_block() {
    const char *e = std::getenv("ASTRA_AMD_FP3D_ANGLE_BLOCK");
    const long x = e ? std::strtol(e, nullptr, 10) : 0;
    return (x >= 1 && x <= 64) ? (unsigned)x : 64u;
}

// Split the angles [first, first + count) of a group's list over the 8
// XCDs in equal counts: xfirst[x] is the first angle of XCD x (relative to
// `first`), xfirst[8] = count.  (Splitting by an estimated cost per angle
// -- samples, or cache lines fetched -- measured slower at 512^3 / 720 x
// 768^2: 44.5-45.9 against 43.8-44.3 ms per group launch; the per-angle time
// varies less with the angle than either estimate.)
void xcd_split(unsigned count, unsigned (&xfirst)[9]) {
    for (unsigned x = 0; x <= 8; ++x) xfirst[x] = (unsigned)((size_t)count * x / 8);
}

// Launch `kern` over the angles of one direction group, in as many
// launches as the 2^32 work-item limit of one grid requires.
template <typename K>
void fp_group_launches(FPArgs A, unsigned count, K kern, hipStream_t s) {
    A.tiles = ((A.nu + 15) / 16) * ((A.nv + 15) / 16);
    A.ablk = fp_angle_block();
    const size_t max_groups = (1u << 24) - 8;  // 256 threads each
    AMX_REQUIRE((size_t)A.tiles * 16 <= max_groups, "fp3d: detector too large for one launch");
    // at most max_groups / tiles - 8 angles per launch: 8 XCD ranges of
    // ceil(count / 8) angles fit
    const unsigned per_launch = (unsigned)(max_groups / A.tiles - 8);
    for (unsigned first = 0; first < count; first += per_launch) {
        A.lfirst = first;
        A.lcount = std::min(count - first, per_launch);
        xcd_split(A.lcount, A.xfirst);
        const size_t groups = (size_t)8 * ((A.lcount + 7) / 8) * A.tiles;
        kern(A, (unsigned)groups);
        AMX_LAUNCHED();
    }
}

// One resident-grid FP launch over `groups` logical workgroups (8 x the
// items per XCD): clear the ticket counters (a kernel, so the launch can be
// captured), then as many workgroups as are resident at once (a multiple
// of 8, from the kernel's occupancy).
void fp_resident_launch(void (*kern)(FPArgs, unsigned *, unsigned), const FPArgs &a,
                        unsigned *tick, unsigned groups, hipStream_t s) {
    static std::mutex mu;
    static std::map<std::pair<void *, int>, unsigned> grids;  // (kernel, device) -> grid
    int dev = 0;
    AMX_HIP(hipGetDevice(&dev));
    unsigned grid;
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = grids.find({(void *)kern, dev});
        if (it == grids.end()) {
            int per_cu = 0, cus = 0;
            AMX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 256, 0));
            AMX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
            it = grids.emplace(std::make_pair((void *)kern, dev),
                               std::max(8u, (unsigned)(std::max(per_cu, 1) * std::max(cus, 1)) & ~7u))
                     .first;
        }
        grid = it->second;
    }
    fp_tick_reset<<<1, 64, 0, s>>>(tick);
    AMX_LAUNCHED();
    kern<<<grid, 256, 0, s>>>(a, tick, groups / 8);
}

// ASTRA_AMD_FP3D_LDS=1: the FP stages volume slices in LDS for groups of
// kFpAngles angles (read at every call)
bool fp_lds() {
    const char *e = std::getenv("ASTRA_AMD_FP3D_LDS");
    return e && e[0] == '1';
}

// fp3d_lds_kernel over the angles of one direction group: one workgroup per
// (detector tile, kFpAngles angles), blocks of 16 angles (fp_lds_map)
void fp_lds_launches(FPArgs A, unsigned count, bool cone, hipStream_t s) {
    A.tiles = ((A.nu + 15) / 16) * ((A.nv + 15) / 16);
    A.ablk = 16;
    const size_t per_block = (size_t)A.tiles * (16 / kFpAngles);
    const size_t max_groups = (1u << 24) - 8;
    AMX_REQUIRE(per_block <= max_groups, "fp3d: detector too large for one launch");
    const unsigned blocks_per_launch = (unsigned)(max_groups / per_block);
    for (unsigned first = 0; first < count; first += blocks_per_launch * 16) {
        A.lfirst = first;
        A.lcount = std::min(count - first, blocks_per_launch * 16);
        const unsigned nblk = (A.lcount + 15) / 16;
        const unsigned groups = (unsigned)(((size_t)nblk * per_block + 7) & ~(size_t)7);
        if (cone)
            fp3d_lds_kernel<GEO_CONE><<<groups, 256, 0, s>>>(A);
        else
            fp3d_lds_kernel<GEO_PAR3D><<<groups, 256, 0, s>>>(A);
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
    // the cylindrical-detector FP has no sub-rays (ConeCylFP ignores
    // iRaysPerDetDim, cuda/3d/cone_cyl.cu:613-666)
    if (g.kind == ASTRA_AMD_GEOM_CYLCONE) ss = 1;
    size_t need = 0;
    int groups[3], ng = 0;
    for (int k = 0; k < 3; ++k)
        if (g.dir_count[k]) {
            need = std::max(need, group_elems(g, k));
            groups[ng++] = k;
        }
    Scratch copy(need * sizeof(float2), s);  // row-pair copy
    Scratch tick(8 * sizeof(unsigned), s);    // ticket counters of the resident FP grid
    for (int i = 0; i < ng; ++i) {
        const int k = groups[i];
        make_group_copy(g, k, vol, copy.as<float2>(), s);
        FPArgs A = fp_args(g, k, out_scale);
        // march3d's buffer loads address a slice at 32-bit scalar offsets
        // of up to 3 slices
        AMX_REQUIRE((size_t)(A.n1 + 2 * kPad) * (A.n2 + 2 * kPad) * sizeof(float2) <= (1ull << 30),
                    "fp3d: volume slice too large (%u x %u)", A.n1, A.n2);
        A.volp = copy.as<float>();
        A.out = proj;
        A.rb = resid_b;
        A.rw = resid_w;
        A.padded_out = padded_out ? 1 : 0;
        // The direction groups partition the angles: every output element
        // is written by exactly one group, which applies the full mode.
        A.mode = mode;
        ProfToken pt = prof_begin(g.kind == ASTRA_AMD_GEOM_CONE     ? "fp3d_cone"
                                  : g.kind == ASTRA_AMD_GEOM_CYLCONE ? "fp3d_cyl"
                                                                     : "fp3d_par3d",
                                  (double)g.dir_count[k] * g.nu * g.nv * A.n0 * ss * ss, s);
        A.ss = ss;
        if (ss == 1 && g.kind != ASTRA_AMD_GEOM_CYLCONE && fp_lds()) {
            fp_lds_launches(A, g.dir_count[k], g.kind == ASTRA_AMD_GEOM_CONE, s);
            prof_end(pt, s);
            continue;
        }
        fp_group_launches(A, g.dir_count[k], [&](const FPArgs &a, unsigned groups) {
            // groups = 8 x the items per XCD
            if (g.kind == ASTRA_AMD_GEOM_CYLCONE)
                fp_resident_launch(fp3d_kernel<GEO_CYL, false>, a, tick.as<unsigned>(), groups, s);
            else if (ss > 1) {
                if (g.kind == ASTRA_AMD_GEOM_CONE)
                    fp_resident_launch(fp3d_kernel<GEO_CONE, true>, a, tick.as<unsigned>(), groups, s);
                else
                    fp_resident_launch(fp3d_kernel<GEO_PAR3D, true>, a, tick.as<unsigned>(), groups,
                                       s);
            } else {
                if (g.kind == ASTRA_AMD_GEOM_CONE)
                    fp_resident_launch(fp3d_kernel<GEO_CONE, false>, a, tick.as<unsigned>(), groups,
                                       s);
                else
                    fp_resident_launch(fp3d_kernel<GEO_PAR3D, false>, a, tick.as<unsigned>(), groups,
                                       s);
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
template <int ZS>
void bp3d_dispatch(int kind, bool zc, const BPArgs &A, hipStream_t s) {
    dim3 grid(((A.nx + 15) / 16) * ((A.ny + 15) / 16), (A.nz + ZS - 1) / ZS);
    if (zc && bp_lds()) {
        const int pm = bp_lds_pitch_mod();
        if (kind == 0) bp3d_lds_kernel<0, ZS><<<grid, 256, 0, s>>>(A, pm);
        else if (kind == 1) bp3d_lds_kernel<1, ZS><<<grid, 256, 0, s>>>(A, pm);
        else bp3d_lds_kernel<2, ZS><<<grid, 256, 0, s>>>(A, pm);
        return;
    }
    if (kind == 0)
        zc ? bp3d_kernel<0, ZS, true><<<grid, 256, 0, s>>>(A) : bp3d_kernel<0, ZS, false><<<grid, 256, 0, s>>>(A);
    else if (kind == 1)
        zc ? bp3d_kernel<1, ZS, true><<<grid, 256, 0, s>>>(A) : bp3d_kernel<1, ZS, false><<<grid, 256, 0, s>>>(A);
    else
        zc ? bp3d_kernel<2, ZS, true><<<grid, 256, 0, s>>>(A) : bp3d_kernel<2, ZS, false><<<grid, 256, 0, s>>>(A);
}

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
    if (g.kind == ASTRA_AMD_GEOM_CYLCONE) {
        // default kernel only, no voxel supersampling (ConeCylBP_Array,
        // cuda/3d/cone_cyl_bp.cu:234-296; cuda/3d/mem3d.cu:463-479)
        if (kernel != ASTRA_AMD_KER3D_DEFAULT)
            fail(ASTRA_AMD_ERR_UNSUPPORTED,
                 "bp3d: the cylindrical-detector geometry supports the default and matched_bp "
                 "kernels only");
        ProfToken pt = prof_begin("bp3d_cyl", (double)g.nx * g.ny * g.nz * g.na, s);
        bp3d_cyl_kernel<ZS><<<grid, 256, 0, s>>>(A, g.bp_cyl.as<CylBPAngle>(), vx, vy, vz);
        AMX_LAUNCHED();
        prof_end(pt, s);
        return;
    }
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
    } else {
        const int zs = bp_zs();
        const size_t plane = (size_t)(A.nv + 2 * kPad) * (A.nu + 2 * kPad) * 2;
        const unsigned na = A.na;
        const unsigned chunk = bp_achunk(plane * na * sizeof(float), na);
        // SET / ADD chunks accumulate in the volume itself; an update mode
        // (SIRT, EM) needs the whole BP first, so the earlier chunks sum into
        // a scratch volume that the last chunk adds before its epilogue.
        const bool direct = mode == OUT_SET || mode == OUT_ADD;
        std::unique_ptr<Scratch> part;
        if (!direct && chunk < na) part.reset(new Scratch((size_t)A.nx * A.ny * A.nz * sizeof(float), s));
        BPArgs C = A;
        for (unsigned first = 0; first == 0 || first < na; first += chunk) {
            const bool last = first + chunk >= na;
            C.na = std::min(chunk, na - first);
            C.ang = A.ang + first;
            C.aw = A.aw ? A.aw + first : nullptr;
            C.pp = A.pp + first * plane;
            if (direct) {
                C.mode = first == 0 ? A.mode : OUT_ADD;
            } else if (!last) {
                C.vol = part->as<float>();
                C.mode = first == 0 ? OUT_SET : OUT_ADD;
            } else {
                C.vol = A.vol;
                C.mode = A.mode;
                C.pre = part ? part->as<float>() : nullptr;
            }
            if (zs == 4) bp3d_dispatch<4>(kind, zc, C, s);
            else if (zs == 16) bp3d_dispatch<16>(kind, zc, C, s);
            else bp3d_dispatch<8>(kind, zc, C, s);
        }
    }
    AMX_LAUNCHED();
    prof_end(pt, s);
}

}  // namespace

void bp3d_launch_padded(const Geom3D &g, const float *pp, float *vol, float out_scale, int kernel,
                        int mode, const float *upd_w, float lo, float hi, hipStream_t s,
                        unsigned ss) {
    Scratch pairs(padded_am_elems(g) * sizeof(float2), s);
    am_to_pairs<<<g.na * ((g.nv + 2 * kPad + kPairRows - 1) / kPairRows), 256, 0, s>>>(
        pp, pairs.as<float2>(), g.nu, g.nv);
    AMX_LAUNCHED();
    bp3d_run_pairs(g, pairs.as<float2>(), vol, out_scale, kernel, mode, upd_w, lo, hi, s, ss);
}

void bp3d_launch(const Geom3D &g, const float *proj, float *vol, float out_scale, int kernel,
                 int mode, const float *upd_w, float lo, float hi, hipStream_t s, unsigned ss) {
    Scratch pairs(padded_am_elems(g) * sizeof(float2), s);
    proj_to_am_pairs<<<g.na * ((g.nv + 2 * kPad + kPairRows - 1) / kPairRows), 256, 0, s>>>(
        proj, pairs.as<float2>(), g.nu, g.na, g.nv);
    AMX_LAUNCHED();
    bp3d_run_pairs(g, pairs.as<float2>(), vol, out_scale, kernel, mode, upd_w, lo, hi, s, ss);
}

// slices per thread of the matched gather (ASTRA_AMD_MATCHED_SLICES:
// 1, 2, 4, 8 or 16; default 8)
static int matched_slices() {
    static const int v = [] {
        const char *e = std::getenv("ASTRA_AMD_MATCHED_SLICES");
        const int x = e ? std::atoi(e) : 8;
        return (x == 1 || x == 2 || x == 4 || x == 16) ? x : 8;
    }();
    return v;
}

void bp3d_matched_launch(const Geom3D &g, const float *proj, float *vol, float out_scale,
                         int mode, hipStream_t s, unsigned ss) {
    AMX_REQUIRE(ss >= 1 && ss <= 16, "bp3d: rays per detector dimension must be in [1, 16]");
    // the transpose of the FP actually taken (no sub-rays on a cylinder)
    if (g.kind == ASTRA_AMD_GEOM_CYLCONE) ss = 1;
    const size_t nvox = (size_t)g.nx * g.ny * g.nz;
    if (mode == OUT_SET) fill_launch(vol, nvox, 0.f, s);
    const size_t NU = (size_t)g.nu * ss, NV = (size_t)g.nv * ss;
    // 32-bit byte offsets into one angle's ray table
    AMX_REQUIRE(NU * NV * sizeof(float4) < (1ull << 32),
                "bp3d: detector too large for the matched adjoint");
    // ray tables are built per chunk of angles, at most 128 MiB at a time
    // (one angle at least): the tables are scratch of every gradient step
    // of the differentiable projector, so they stay small beside the data
    const size_t per_angle = NU * NV * (sizeof(float4) + sizeof(float));
    const unsigned chunk = (unsigned)std::max<size_t>(
        1, std::min<size_t>(g.na, (size_t(128) << 20) / per_angle));
    const bool cyl = g.kind == ASTRA_AMD_GEOM_CYLCONE;
    Scratch rays(per_angle * chunk, s), geo((size_t)chunk * 3 * sizeof(float4), s);
    float4 *rp = rays.as<float4>();
    float *yp = reinterpret_cast<float *>(rp + NU * NV * chunk);
    for (int k = 0; k < 3; ++k) {
        if (!g.dir_count[k]) continue;
        FPArgs A = fp_args(g, k, out_scale);
        A.ss = ss;
        MatchedArgs M{};
        M.rays = rp;
        M.yw = yp;
        M.geo = geo.as<float4>();
        M.NU = (unsigned)NU;
        M.NV = (unsigned)NV;
        M.n0 = A.n0; M.n1 = A.n1; M.n2 = A.n2;
        M.tiles1 = (A.n1 + 15) / 16;
        M.nx = g.nx; M.ny = g.ny;
        M.group = k;
        M.vol = vol;
        M.cyl = g.bp_cyl.as<CylBPAngle>();
        for (int q = 0; q < 3; ++q) M.vs[q] = g.vol_scale[q];
        const int S = cyl ? 8 : matched_slices();
        const size_t gblocks = (size_t)((A.n0 + S - 1) / S) * M.tiles1 * ((A.n2 + 15) / 16);
        AMX_REQUIRE(gblocks < (1ull << 31), "bp3d: volume too large for one matched launch");
        for (unsigned first = 0; first < (unsigned)g.dir_count[k]; first += chunk) {
            A.lfirst = first;
            A.lcount = std::min((unsigned)g.dir_count[k] - first, chunk);
            const size_t nrays = NU * NV * A.lcount;
            const unsigned rblocks = (unsigned)std::min<size_t>((nrays + 255) / 256, 1u << 20);
            const unsigned gb = (A.lcount + 63) / 64;
            if (cyl) {
                matched_rays_kernel<GEO_CYL><<<rblocks, 256, 0, s>>>(A, proj, rp, yp);
            } else if (g.kind == ASTRA_AMD_GEOM_CONE) {
                matched_rays_kernel<GEO_CONE><<<rblocks, 256, 0, s>>>(A, proj, rp, yp);
                matched_geo_kernel<true><<<gb, 64, 0, s>>>(A, geo.as<float4>());
            } else {
                matched_rays_kernel<GEO_PAR3D><<<rblocks, 256, 0, s>>>(A, proj, rp, yp);
                matched_geo_kernel<false><<<gb, 64, 0, s>>>(A, geo.as<float4>());
            }
            M.count = A.lcount;
            M.alist = A.alist + first;
            const unsigned gbl = (unsigned)gblocks;
            if (cyl)
                bp3d_matched_gather<8, true><<<gbl, 256, 0, s>>>(M);
            else if (S == 1)
                bp3d_matched_gather<1, false><<<gbl, 256, 0, s>>>(M);
            else if (S == 2)
                bp3d_matched_gather<2, false><<<gbl, 256, 0, s>>>(M);
            else if (S == 4)
                bp3d_matched_gather<4, false><<<gbl, 256, 0, s>>>(M);
            else if (S == 8)
                bp3d_matched_gather<8, false><<<gbl, 256, 0, s>>>(M);
            else
                bp3d_matched_gather<16, false><<<gbl, 256, 0, s>>>(M);
            AMX_LAUNCHED();
        }
    }
}

}  // namespace amx
