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
    for (unsigned len = 2, step = half_n; len <= N; len <<= 1, step >>= 1) {
        const unsigned h = len >> 1;
        for (unsigned b = threadIdx.x; b < half_n; b += blockDim.x) {
            unsigned k = b & (h - 1);
            unsigned i = (b - k) * 2 + k;
            float2 t = cmul(X[i + h], A.tw[k * step]);
            float2 u = X[i];
            X[i] = make_float2(u.x + t.x, u.y + t.y);
            X[i + h] = make_float2(u.x - t.x, u.y - t.y);
        }
        __syncthreads();
    }

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
    for (unsigned len = N, step = 1; len >= 2; len >>= 1, step <<= 1) {
        const unsigned h = len >> 1;
        for (unsigned b = threadIdx.x; b < half_n; b += blockDim.x) {
            unsigned k = b & (h - 1);
            unsigned i = (b - k) * 2 + k;
            float2 u = X[i], v = X[i + h];
            X[i] = make_float2(u.x + v.x, u.y + v.y);
            X[i + h] = cmul(make_float2(u.x - v.x, u.y - v.y), A.tw[k * step]);
        }
        __syncthreads();
    }

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
    for (unsigned len = 2, step = M; len <= M; len <<= 1, step >>= 1) {
        const unsigned h = len >> 1;
        for (unsigned b = threadIdx.x; b < half_m; b += blockDim.x) {
            unsigned k = b & (h - 1);
            unsigned i = (b - k) * 2 + k;
            float2 t = cmul(Z[i + h], A.tw[k * step]);
            float2 u = Z[i];
            Z[i] = make_float2(u.x + t.x, u.y + t.y);
            Z[i + h] = make_float2(u.x - t.x, u.y - t.y);
        }
        __syncthreads();
    }

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
    for (unsigned len = M, step = 2; len >= 2; len >>= 1, step <<= 1) {
        const unsigned h = len >> 1;
        for (unsigned b = threadIdx.x; b < half_m; b += blockDim.x) {
            unsigned k = b & (h - 1);
            unsigned i = (b - k) * 2 + k;
            float2 u = Z[i], q = Z[i + h];
            Z[i] = make_float2(u.x + q.x, u.y + q.y);
            Z[i + h] = cmul(make_float2(u.x - q.x, u.y - q.y), A.tw[k * step]);
        }
        __syncthreads();
    }

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
