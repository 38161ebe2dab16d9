// Row filtering for FBP / FDK on gfx950: a hand-written batched radix-2 FFT
// in LDS.  Two real rows share one complex transform (row a in the real
// part, row a+1 in the imaginary part; rows too long for a pair in LDS
// take filter_row_real_kernel, one row per workgroup): with a real, even filter h the
// filtered rows are Re and Im of IFFT(H . FFT(x + i y)).  Forward pass is
// decimation-in-time (bit-reversed load, natural-order spectrum), the
// inverse is decimation-in-frequency (natural input, bit-reversed output),
// so no explicit reordering pass is needed.  The optional FDK cosine
// pre-weighting is applied while loading (one HBM read + one HBM write per
// sample for pre-weight + filter together).
//
// Semantics: zero-pad each row of n_det samples to n_fft, multiply the
// spectrum by the real filter, inverse transform, divide by n_fft and keep
// the first n_det samples (cuda/3d/fdk.cu:351-398, cuda/2d/fft.cu:52-131).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>

#include "devutil.hpp"
#include "internal.hpp"

namespace amx {

struct FilterArgs {
    float *rows;
    size_t n_rows;
    unsigned n_det, n_fft, log2n;
    const float *spec;    // [n_fft] real spectrum (shared by every row)
    const float2 *cspec;  // general: [n_spec][n_fft] complex spectra
    unsigned n_spec;      // general: row r uses spectrum r % n_spec
    const float2 *tw;     // [n_fft/2] exp(-2 pi i k / n_fft)
    const float *weights; // optional [n_rows / rows_per_weight][n_det]
    size_t rows_per_weight;
    const float *weights2; // optional second factor [w2_rows][n_det]: row r uses r % w2_rows
    unsigned w2_rows;
};

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ __forceinline__ float2 conjf2(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 addf2(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 subf2(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }

// In-place FFT over n points in LDS: decimation in time, bit-reversed input,
// natural-order output.  The radix-2 stages are taken two at a time (a
// radix-2^2 butterfly on 4 points held in registers: the same operations
// as two radix-2 passes, with half the LDS round trips and barriers); an
// odd stage count ends with one radix-2 stage.  The twiddle of position k
// in a stage of length len is tw[k * (n / len) * ts], where the table has
// length T = n * ts (tw[j] = exp(-2 pi i j / T), j < T / 2).
__device__ __forceinline__ void lds_fft_dit(float2 *X, unsigned n, const float2 *__restrict__ tw,
                                            unsigned ts) {
    unsigned len = 2;
    for (; 2 * len <= n; len <<= 2) {
        const unsigned h = len >> 1, s1 = (n / len) * ts, s2 = s1 >> 1;
        for (unsigned g = threadIdx.x; g < (n >> 2); g += blockDim.x) {
            const unsigned k = g & (h - 1), a = (g - k) * 4 + k;
            float2 xa = X[a], xb = X[a + h], xc = X[a + 2 * h], xd = X[a + 3 * h];
            const float2 w1 = tw[k * s1];
            float2 t = cmul(xb, w1);
            xb = subf2(xa, t);
            xa = addf2(xa, t);
            t = cmul(xd, w1);
            xd = subf2(xc, t);
            xc = addf2(xc, t);
            t = cmul(xc, tw[k * s2]);
            X[a + 2 * h] = subf2(xa, t);
            X[a] = addf2(xa, t);
            t = cmul(xd, tw[(k + h) * s2]);
            X[a + 3 * h] = subf2(xb, t);
            X[a + h] = addf2(xb, t);
        }
        __syncthreads();
    }
    if (len <= n) {
        const unsigned h = len >> 1, s1 = (n / len) * ts;
        for (unsigned b = threadIdx.x; b < (n >> 1); b += blockDim.x) {
            const unsigned k = b & (h - 1), i = (b - k) * 2 + k;
            const float2 t = cmul(X[i + h], tw[k * s1]);
            const float2 u = X[i];
            X[i] = addf2(u, t);
            X[i + h] = subf2(u, t);
        }
        __syncthreads();
    }
}

// In-place FFT over n points in LDS: decimation in frequency, natural
// input, bit-reversed output; stages two at a time as in lds_fft_dit.
__device__ __forceinline__ void lds_fft_dif(float2 *X, unsigned n, const float2 *__restrict__ tw,
                                            unsigned ts) {
    unsigned len = n;
    for (; len >= 4; len >>= 2) {
        const unsigned q = len >> 2, s1 = (n / len) * ts, s2 = s1 << 1;
        for (unsigned g = threadIdx.x; g < (n >> 2); g += blockDim.x) {
            const unsigned k = g & (q - 1), a = (g - k) * 4 + k;
            const float2 xa = X[a], xb = X[a + q], xc = X[a + 2 * q], xd = X[a + 3 * q];
            const float2 ya = addf2(xa, xc), yc = cmul(subf2(xa, xc), tw[k * s1]);
            const float2 yb = addf2(xb, xd), yd = cmul(subf2(xb, xd), tw[(k + q) * s1]);
            const float2 w2 = tw[k * s2];
            X[a] = addf2(ya, yb);
            X[a + q] = cmul(subf2(ya, yb), w2);
            X[a + 2 * q] = addf2(yc, yd);
            X[a + 3 * q] = cmul(subf2(yc, yd), w2);
        }
        __syncthreads();
    }
    if (len == 2) {
        for (unsigned b = threadIdx.x; b < (n >> 1); b += blockDim.x) {
            const unsigned i = 2 * b;
            const float2 u = X[i], v = X[i + 1];
            X[i] = addf2(u, v);
            X[i + 1] = cmul(subf2(u, v), tw[0]);
        }
        __syncthreads();
    }
}

// GENERAL = false: one real, even spectrum for both rows (the pair's
// transform is simply scaled).  GENERAL = true: each row has its own
// (complex, Hermitian) spectrum; the two rows' spectra are separated from
// the shared transform Z as X[k] = (Z[k] + conj Z[N-k]) / 2 and
// Y[k] = (Z[k] - conj Z[N-k]) / 2i, filtered, and recombined as
// W = Ha X + i Hb Y, whose inverse holds the filtered rows in its real and
// imaginary parts.
template <bool GENERAL>
__global__ __launch_bounds__(256) void filter_rows_kernel(FilterArgs A) {
    extern __shared__ float2 X[];
    const unsigned N = A.n_fft;
    const unsigned half_n = N >> 1;
    const size_t r0 = (size_t)blockIdx.x * 2;
    const bool has1 = r0 + 1 < A.n_rows;
    float *row0 = A.rows + r0 * A.n_det;
    float *row1 = row0 + A.n_det;
    const float *w0 = A.weights ? A.weights + (r0 / A.rows_per_weight) * A.n_det : nullptr;
    const float *w1 = (A.weights && has1) ? A.weights + ((r0 + 1) / A.rows_per_weight) * A.n_det : nullptr;
    const float *v0 = A.weights2 ? A.weights2 + (r0 % A.w2_rows) * A.n_det : nullptr;
    const float *v1 = (A.weights2 && has1) ? A.weights2 + ((r0 + 1) % A.w2_rows) * A.n_det : nullptr;
    const unsigned shift = 32 - A.log2n;

    // load (bit-reversed positions), zero padding beyond n_det
    for (unsigned j = threadIdx.x; j < N; j += blockDim.x) {
        float a = 0.f, b = 0.f;
        if (j < A.n_det) {
            a = row0[j];
            if (w0) a *= w0[j];
            if (v0) a *= v0[j];
            if (has1) {
                b = row1[j];
                if (w1) b *= w1[j];
                if (v1) b *= v1[j];
            }
        }
        X[__brev(j) >> shift] = make_float2(a, b);
    }
    __syncthreads();

    // forward DIT: natural-order spectrum
    lds_fft_dit(X, N, A.tw, 1);

    // x filter, conjugate (inverse via conj(FFT(conj(.))))
    if (!GENERAL) {
        for (unsigned k = threadIdx.x; k < N; k += blockDim.x) {
            float hk = A.spec[k];
            float2 v = X[k];
            X[k] = make_float2(v.x * hk, -v.y * hk);
        }
    } else {
        const float2 *Ha = A.cspec + (size_t)(r0 % A.n_spec) * N;
        const float2 *Hb = has1 ? A.cspec + (size_t)((r0 + 1) % A.n_spec) * N : Ha;
        // pairs (k, N - k) are disjoint for k = 0 .. N/2: one thread each
        for (unsigned k = threadIdx.x; k <= half_n; k += blockDim.x) {
            const unsigned m = (N - k) & (N - 1);
            const float2 z1 = X[k], z2 = X[m];
            const float2 xk = make_float2(0.5f * (z1.x + z2.x), 0.5f * (z1.y - z2.y));
            const float2 d = make_float2(0.5f * (z1.x - z2.x), 0.5f * (z1.y + z2.y));
            const float2 yk = make_float2(d.y, -d.x);  // d / i
            const float2 a = cmul(Ha[k], xk), b = cmul(Hb[k], yk);
            X[k] = conjf2(make_float2(a.x - b.y, a.y + b.x));  // conj(a + i b)
            if (m != k) {
                const float2 a2 = cmul(Ha[m], conjf2(xk)), b2 = cmul(Hb[m], conjf2(yk));
                X[m] = conjf2(make_float2(a2.x - b2.y, a2.y + b2.x));
            }
        }
    }
    __syncthreads();

    // DIF: bit-reversed output
    lds_fft_dif(X, N, A.tw, 1);

    // y = conj(F) / N: row0 = Re F / N, row1 = -Im F / N
    const float inv = 1.0f / (float)N;
    for (unsigned j = threadIdx.x; j < A.n_det; j += blockDim.x) {
        float2 v = X[__brev(j) >> shift];
        row0[j] = v.x * inv;
        if (has1) row1[j] = -v.y * inv;
    }
}

// One real row per workgroup, for transforms whose row pair does not fit
// the CU's LDS (n_fft > 16384 on gfx950): the length-N real transform is a
// length-M = N/2 complex transform of z[m] = x[2m] + i x[2m+1] followed by
// the even / odd split
//   E[k] = (Z[k] + conj Z[M-k]) / 2,  O[k] = (Z[k] - conj Z[M-k]) / 2i,
//   X[k] = E[k] + W^k O[k],  X[k+M] = E[k] - W^k O[k],  W = exp(-2 pi i / N),
// and back: with the filtered X' (Hermitian, so X'[k+M] = conj X'[M-k]),
//   E'[k] = (X'[k] + conj X'[M-k]) / 2,  O'[k] = (X'[k] - conj X'[M-k]) / (2 W^k),
// z' = IFFT_M(E' + i O') holds x'[2m] = Re z'[m] and x'[2m+1] = Im z'[m].
// The filter multiplies X[k], k = 0..M (real even spectrum, or row r's
// complex Hermitian spectrum r % n_spec).  M points of float2 in LDS.
template <bool GENERAL>
__global__ __launch_bounds__(256) void filter_row_real_kernel(FilterArgs A) {
    extern __shared__ float2 Z[];
    const unsigned N = A.n_fft, M = N >> 1, half_m = M >> 1;
    const size_t r = blockIdx.x;
    float *row = A.rows + r * A.n_det;
    const float *w = A.weights ? A.weights + (r / A.rows_per_weight) * A.n_det : nullptr;
    const float *v = A.weights2 ? A.weights2 + (r % A.w2_rows) * A.n_det : nullptr;
    const unsigned shift = 32 - (A.log2n - 1);

    // load pairs (x[2m], x[2m+1]) at bit-reversed positions of the M-point transform
    for (unsigned m = threadIdx.x; m < M; m += blockDim.x) {
        float a = 0.f, b = 0.f;
        const unsigned j = 2 * m;
        if (j < A.n_det) {
            a = row[j];
            if (w) a *= w[j];
            if (v) a *= v[j];
        }
        if (j + 1 < A.n_det) {
            b = row[j + 1];
            if (w) b *= w[j + 1];
            if (v) b *= v[j + 1];
        }
        Z[__brev(m) >> shift] = make_float2(a, b);
    }
    __syncthreads();

    // forward DIT over M points; exp(-2 pi i k / len) = tw[k * N / len]
    lds_fft_dit(Z, M, A.tw, 2);

    // split, filter, merge: one thread per pair (k, M - k), k = 0 .. M/2;
    // the conjugate of Z' is stored for the inverse (conj(FFT(conj .)))
    const float2 *H = GENERAL ? A.cspec + (size_t)(r % A.n_spec) * N : nullptr;
    for (unsigned k = threadIdx.x; k <= half_m; k += blockDim.x) {
        const unsigned j = M - k;             // == M for k == 0
        const float2 zk = Z[k], zj = Z[j & (M - 1)];
        // E[k], O[k] and the spectrum at k and at j (for k == 0: X[0], X[M])
        const float2 ek = make_float2(0.5f * (zk.x + zj.x), 0.5f * (zk.y - zj.y));
        const float2 dk = make_float2(0.5f * (zk.x - zj.x), 0.5f * (zk.y + zj.y));
        const float2 ok = make_float2(dk.y, -dk.x);                  // d / i
        const float2 wk = A.tw[k];                                   // W^k, k < M
        const float2 wok = cmul(wk, ok);
        float2 xk = make_float2(ek.x + wok.x, ek.y + wok.y);
        // X[j] = conj E[k] + W^j conj O[k], W^j = -conj W^k
        const float2 wj = make_float2(-wk.x, wk.y);
        const float2 woj = cmul(wj, conjf2(ok));
        float2 xj = make_float2(ek.x + woj.x, -ek.y + woj.y);
        if (GENERAL) {
            xk = cmul(H[k], xk);
            xj = cmul(H[j], xj);
        } else {
            const float hk = A.spec[k], hj = A.spec[j];
            xk = make_float2(xk.x * hk, xk.y * hk);
            xj = make_float2(xj.x * hj, xj.y * hj);
        }
        // E'[k] = (X'[k] + conj X'[j]) / 2, O'[k] = (X'[k] - conj X'[j]) conj(W^k) / 2
        const float2 e2 = make_float2(0.5f * (xk.x + xj.x), 0.5f * (xk.y - xj.y));
        const float2 o2 = cmul(make_float2(0.5f * (xk.x - xj.x), 0.5f * (xk.y + xj.y)), conjf2(wk));
        // Z'[k] = E' + i O'; stored conjugated
        Z[k] = make_float2(e2.x - o2.y, -(e2.y + o2.x));
        if (k != 0 && j != k) {
            // E'[j] = conj E'[k]; O'[j] = (X'[j] - conj X'[k]) conj(W^j) / 2
            const float2 o3 = cmul(make_float2(0.5f * (xj.x - xk.x), 0.5f * (xj.y + xk.y)), conjf2(wj));
            const float2 e3 = conjf2(e2);
            Z[j] = make_float2(e3.x - o3.y, -(e3.y + o3.x));
        }
    }
    __syncthreads();

    // DIF over M points: bit-reversed output
    lds_fft_dif(Z, M, A.tw, 2);

    // z' = conj(result) / M: x'[2m] = Re / M, x'[2m+1] = -Im / M
    const float inv = 1.0f / (float)M;
    for (unsigned m = threadIdx.x; m < M; m += blockDim.x) {
        const unsigned j = 2 * m;
        if (j >= A.n_det) break;
        const float2 q = Z[__brev(m) >> shift];
        row[j] = q.x * inv;
        if (j + 1 < A.n_det) row[j + 1] = -q.y * inv;
    }
}

// rows *= w[r / div] (optional) * w2[r % mod] (optional): the weights of an
// identity ('none') filter
__global__ void k_row_weights(float *rows, unsigned len, const float *w, size_t div,
                              const float *w2, unsigned mod) {
    const size_t r = blockIdx.x;
    float *row = rows + r * len;
    const float *a = w ? w + (r / div) * len : nullptr;
    const float *b = w2 ? w2 + (r % mod) * len : nullptr;
    for (unsigned i = threadIdx.x; i < len; i += blockDim.x) {
        float v = row[i];
        if (a) v *= a[i];
        if (b) v *= b[i];
        row[i] = v;
    }
}

void filter_rows_launch(const Filter &f, float *rows, size_t n_rows, const float *weights,
                        size_t rows_per_weight, hipStream_t s, const float *weights2,
                        unsigned w2_rows) {
    if (n_rows == 0) return;
    if (!rows_per_weight) rows_per_weight = 1;
    if (!w2_rows) w2_rows = 1;
    if (f.identity) {  // no filtering; the weights still apply
        if (weights || weights2) {
            AMX_REQUIRE(n_rows < (1ull << 31), "filter: too many rows");
            k_row_weights<<<(unsigned)n_rows, 256, 0, s>>>(rows, f.n_det, weights, rows_per_weight,
                                                           weights2, w2_rows);
            AMX_LAUNCHED();
        }
        return;
    }
    FilterArgs A{};
    A.rows = rows;
    A.n_rows = n_rows;
    A.n_det = f.n_det;
    A.n_fft = f.n_fft;
    unsigned l = 0;
    while ((1u << l) < f.n_fft) ++l;
    A.log2n = l;
    A.spec = f.general ? nullptr : f.spec.as<float>();
    A.cspec = f.general ? f.spec.as<float2>() : nullptr;
    A.n_spec = f.n_spec;
    A.tw = f.twiddle.as<float2>();
    A.weights = weights;
    A.rows_per_weight = rows_per_weight;
    A.weights2 = weights2;
    A.w2_rows = w2_rows;
    // One row pair's transform lives in LDS: n_fft <= 8192 fits the default
    // 64 KiB; larger transforms (up to the CU's LDS, 160 KiB on gfx950)
    // raise the kernel's dynamic LDS limit first.  Beyond that, one real
    // row per workgroup as a half-length complex transform.
    int dev = 0, max_lds = 0;
    AMX_HIP(hipGetDevice(&dev));
    AMX_HIP(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev));
    const size_t pair_lds = (size_t)f.n_fft * sizeof(float2);
    // ASTRA_AMD_FILTER_SINGLE_ROW=1 takes the single-row path at any size
    // (read at every call; for testing it where the pair path also fits)
    const char *single = std::getenv("ASTRA_AMD_FILTER_SINGLE_ROW");
    if (pair_lds > (size_t)max_lds || (single && single[0] == '1')) {
        const size_t lds = pair_lds / 2;
        AMX_REQUIRE(f.n_fft >= 4 && lds <= (size_t)max_lds,
                    "filter: n_fft %u needs %zu bytes of LDS for one row, the device has %d", f.n_fft,
                    lds, max_lds);
        if (f.general)
            AMX_HIP(hipFuncSetAttribute((const void *)filter_row_real_kernel<true>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        else
            AMX_HIP(hipFuncSetAttribute((const void *)filter_row_real_kernel<false>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        AMX_REQUIRE(n_rows < (1ull << 31), "filter: too many rows");
        ProfToken pt = prof_begin("filter_rows", (double)n_rows, s);
        if (f.general)
            filter_row_real_kernel<true><<<(unsigned)n_rows, 256, lds, s>>>(A);
        else
            filter_row_real_kernel<false><<<(unsigned)n_rows, 256, lds, s>>>(A);
        AMX_LAUNCHED();
        prof_end(pt, s);
        return;
    }
    const size_t lds = pair_lds;
    if (lds > 65536) {
        if (f.general)
            AMX_HIP(hipFuncSetAttribute((const void *)filter_rows_kernel<true>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        else
            AMX_HIP(hipFuncSetAttribute((const void *)filter_rows_kernel<false>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    }
    const size_t groups = (n_rows + 1) / 2;
    AMX_REQUIRE(groups < (1ull << 31), "filter: too many rows");
    ProfToken pt = prof_begin("filter_rows", (double)n_rows, s);
    if (f.general)
        filter_rows_kernel<true><<<(unsigned)groups, 256, lds, s>>>(A);
    else
        filter_rows_kernel<false><<<(unsigned)groups, 256, lds, s>>>(A);
    AMX_LAUNCHED();
    prof_end(pt, s);
}

void fdk_parker(const Geom3D &g, std::vector<float> &w) {
    // angles of FDK() (cuda/3d/fdk.cu:447-450): -atan2(srcX, srcY) + pi, and
    // the scalar geometry of angle 0 (cuda/3d/fdk.cu:427-445)
    const double *p = g.vecs.data();
    std::vector<double> ang(g.na);
    for (unsigned a = 0; a < g.na; ++a)
        ang[a] = -std::atan2(p[12 * a + 0], p[12 * a + 1]) + M_PI;
    const double dcx = p[3] + 0.5 * g.nu * p[6], dcy = p[4] + 0.5 * g.nu * p[7];
    const double sod = std::sqrt(p[0] * p[0] + p[1] * p[1]);
    const double odd = std::sqrt(dcx * dcx + dcy * dcy);
    const double du = std::sqrt(p[6] * p[6] + p[7] * p[7]);
    parker_weights(ang, g.nu, sod, odd, du, w);
}

void fdk_weights(const Geom3D &g, std::vector<float> &w) {
    // devFDK_preweight (cuda/3d/fdk.cu:70-113) with the scalar geometry that
    // FDK() extracts from angle 0 (cuda/3d/fdk.cu:427-445), evaluated in FP64.
    const double *p = g.vecs.data();
    double dcx = p[3] + 0.5 * g.nu * p[6];
    double dcy = p[4] + 0.5 * g.nu * p[7];
    double dcz = p[5] + 0.5 * g.nv * p[11];
    double sod = std::sqrt(p[0] * p[0] + p[1] * p[1]);
    double odd = std::sqrt(dcx * dcx + dcy * dcy);
    double du = std::sqrt(p[6] * p[6] + p[7] * p[7]);
    double dv = std::fabs(p[11]);
    double zshift = dcz - p[2];
    double D = sod + odd;
    double W = D * (D / (du * sod)) * (M_PI / 2.0) / (double)g.na;
    w.resize((size_t)g.nv * g.nu);
    for (unsigned iv = 0; iv < g.nv; ++iv) {
        double V = (iv - 0.5 * g.nv + 0.5) * dv + zshift;
        for (unsigned iu = 0; iu < g.nu; ++iu) {
            double U = (iu - 0.5 * g.nu + 0.5) * du;
            w[(size_t)iv * g.nu + iu] = (float)(W / std::sqrt(D * D + U * U + V * V));
        }
    }
}

}  // namespace amx
