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

    // load (bit-reversed positions), zero padding beyond n_det; the loads
    // of an element come before its products, so they are in flight
    // together
    for (unsigned j = threadIdx.x; j < N; j += blockDim.x) {
        float a = 0.f, b = 0.f, wa = 1.f, wb = 1.f, va = 1.f, vb = 1.f;
        if (j < A.n_det) {
            a = row0[j];
            if (w0) wa = w0[j];
            if (v0) va = v0[j];
            if (has1) {
                b = row1[j];
                if (w1) wb = w1[j];
                if (v1) vb = v1[j];
            }
        }
        // a product with 1 is exact: the same floats as multiplying only
        // where a weight exists
        X[__brev(j) >> shift] = make_float2(a * wa * va, b * wb * vb);
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

// ---- Stockham FFT in registers (the real, even spectrum of every FDK /
// FBP window) ----------------------------------------------------------
// Each transform of N = 2^LOGN points is held by T = N / 16 threads, 16
// complex values per thread: thread t holds elements t + T m (m = 0..15).
// The passes are radix 16 (the last one radix 16, 8, 4 or 2), each a
// register-resident DFT of its butterflies; between passes the values go
// through LDS once (Stockham ordering: natural-order input and output, no
// bit reversal).  After the forward transform every thread holds its 16
// spectrum values in natural order, so the filter multiply and the first
// pass of the inverse need no exchange.  Pass with radix R after earlier
// radices of product Ns: butterfly j (of N / R) takes in[j + k N / R],
// k < R, multiplies element k by exp(-2 pi i (j mod Ns) k / (Ns R)), and
// its DFT output k goes to (j / Ns) Ns R + (j mod Ns) + k Ns.
// A thread with 16 / R = B butterflies takes j = t + T b (b < B), element k
// of butterfly b being its value m = b + B k, so the thread's inputs are
// always in[t + T m] and the last pass leaves out[t + T m] in value m.
template <int LOGN, unsigned WGT = 256>
struct FftPlan {
    static constexpr unsigned N = 1u << LOGN;
    static constexpr unsigned T = N / 16;                       // threads per transform
    static constexpr unsigned P = (LOGN + 3) / 4;               // passes
    static constexpr unsigned LAST = 1u << (LOGN - 4 * (P - 1));
    static constexpr unsigned F = T >= WGT ? 1 : WGT / T;       // transforms per workgroup
    static constexpr unsigned WG = T * F;
    static constexpr unsigned PADN = N + N / 16;                // LDS float2 per transform
};

// LDS index with one float2 of padding per 16: the 16 lanes that write a
// stride-16 pattern (first pass) land in distinct banks.
__device__ __forceinline__ unsigned padi(unsigned a) { return a + (a >> 4); }

// x * exp(-2 pi i e / 16), e constant after unrolling
template <int E>
__device__ __forceinline__ float2 w16(float2 x) {
    constexpr int e = E & 15;
    if constexpr (e == 0) return x;
    else if constexpr (e == 4) return make_float2(x.y, -x.x);   // -i
    else if constexpr (e == 8) return make_float2(-x.x, -x.y);
    else if constexpr (e == 12) return make_float2(-x.y, x.x);  // +i
    else {
        // cos / sin of 2 pi e / 16 from the first quadrant's values
        constexpr double q[5] = {1.0, 0.92387953251128675613, 0.70710678118654752440,
                                 0.38268343236508977173, 0.0};
        constexpr int qe = e & 3, quad = e >> 2;
        constexpr double c0 = q[qe], s0 = q[4 - qe];
        constexpr float c = (float)(quad == 0 ? c0 : quad == 1 ? -s0 : quad == 2 ? -c0 : s0);
        constexpr float sn = (float)(quad == 0 ? s0 : quad == 1 ? c0 : quad == 2 ? -s0 : -c0);
        // (x.x + i x.y)(c - i sn)
        return make_float2(fmaf(x.x, c, x.y * sn), fmaf(x.y, c, -x.x * sn));
    }
}

__device__ __forceinline__ void dft2(float2 &a, float2 &b) {
    const float2 t = a;
    a = addf2(t, b);
    b = subf2(t, b);
}

__device__ __forceinline__ void dft4(float2 &x0, float2 &x1, float2 &x2, float2 &x3) {
    const float2 a0 = addf2(x0, x2), a1 = subf2(x0, x2), a2 = addf2(x1, x3), a3 = subf2(x1, x3);
    const float2 m3 = make_float2(a3.y, -a3.x);  // -i a3
    x0 = addf2(a0, a2);
    x2 = subf2(a0, a2);
    x1 = addf2(a1, m3);
    x3 = subf2(a1, m3);
}

// In-register forward DFT of R points, natural order in and out:
// X[k] = sum_n x[n] exp(-2 pi i n k / R).  R = 8, 16 as 2 x 4 and 4 x 4
// (inner DFTs over n1 of x[(R / N1')...], twiddles, outer DFTs).
template <unsigned R>
__device__ __forceinline__ void dft(float2 (&x)[R]) {
    if constexpr (R == 2) {
        dft2(x[0], x[1]);
    } else if constexpr (R == 4) {
        dft4(x[0], x[1], x[2], x[3]);
    } else if constexpr (R == 8) {
        // n = 4 n1 + n2, k = k1 + 2 k2: inner radix 2 over n1, twiddle
        // W8^(n2 k1) = W16^(2 n2 k1), radix 4 over n2
#pragma unroll
        for (int n2 = 0; n2 < 4; ++n2) dft2(x[n2], x[4 + n2]);
        x[5] = w16<2>(x[5]);
        x[6] = w16<4>(x[6]);
        x[7] = w16<6>(x[7]);
        dft4(x[0], x[1], x[2], x[3]);  // k1 = 0 -> X[2 k2]
        dft4(x[4], x[5], x[6], x[7]);  // k1 = 1 -> X[1 + 2 k2]
        const float2 y[8] = {x[0], x[4], x[1], x[5], x[2], x[6], x[3], x[7]};
#pragma unroll
        for (int k = 0; k < 8; ++k) x[k] = y[k];
    } else {
        static_assert(R == 16, "radix");
        // n = 4 n1 + n2, k = k1 + 4 k2: inner radix 4 over n1 (x[n2], x[4 +
        // n2], x[8 + n2], x[12 + n2] -> k1 = 0..3), twiddle W16^(n2 k1),
        // radix 4 over n2
#pragma unroll
        for (int n2 = 0; n2 < 4; ++n2) dft4(x[n2], x[4 + n2], x[8 + n2], x[12 + n2]);
        x[5] = w16<1>(x[5]);
        x[6] = w16<2>(x[6]);
        x[7] = w16<3>(x[7]);
        x[9] = w16<2>(x[9]);
        x[10] = w16<4>(x[10]);
        x[11] = w16<6>(x[11]);
        x[13] = w16<3>(x[13]);
        x[14] = w16<6>(x[14]);
        x[15] = w16<9>(x[15]);
        // now x[4 k1 + n2]: radix 4 over n2 for each k1 -> X[k1 + 4 k2]
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) dft4(x[4 * k1], x[4 * k1 + 1], x[4 * k1 + 2], x[4 * k1 + 3]);
        float2 y[16];
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1)
#pragma unroll
            for (int k2 = 0; k2 < 4; ++k2) y[k1 + 4 * k2] = x[4 * k1 + k2];
#pragma unroll
        for (int k = 0; k < 16; ++k) x[k] = y[k];
    }
}

// One pass (radix R, earlier radices of product NS).  Not the last pass:
// the outputs go through LDS (lds = this transform's region) and come back
// as the next pass's inputs in[t + T m].
template <int LOGN, unsigned R, unsigned NS, bool LASTP>
__device__ __forceinline__ void fft_pass(float2 (&v)[16], unsigned t, const float2 *__restrict__ tw,
                                         float2 *lds) {
    using PL = FftPlan<LOGN>;
    constexpr unsigned N = PL::N, T = PL::T, B = 16 / R;
#pragma unroll
    for (unsigned b = 0; b < B; ++b) {
        const unsigned j = t + T * b;
        float2 x[R];
#pragma unroll
        for (unsigned k = 0; k < R; ++k) x[k] = v[b + B * k];
        if constexpr (NS > 1) {
            // w^k for k < R from the table's w, w^2, w^4, w^8 (exact to
            // fp32), the others as products of at most four of them
            const unsigned step = (j & (NS - 1)) * (N / (NS * R));
            float2 w[R];
            w[1] = tw[step];
            if constexpr (R > 2) w[2] = tw[2 * step];
            if constexpr (R > 4) w[4] = tw[4 * step];
            if constexpr (R > 8) w[8] = tw[8 * step];
#pragma unroll
            for (unsigned k = 3; k < R; ++k) {
                const unsigned hb = 1u << (31 - __builtin_clz(k));  // highest power of two in k
                if (k != hb) w[k] = cmul(w[k - hb], w[hb]);
            }
#pragma unroll
            for (unsigned k = 1; k < R; ++k) x[k] = cmul(x[k], w[k]);
        }
        dft<R>(x);
#pragma unroll
        for (unsigned k = 0; k < R; ++k) v[b + B * k] = x[k];
    }
    if constexpr (!LASTP) {
#pragma unroll
        for (unsigned b = 0; b < B; ++b) {
            const unsigned j = t + T * b;
            const unsigned base = (j / NS) * NS * R + (j & (NS - 1));
#pragma unroll
            for (unsigned k = 0; k < R; ++k) lds[padi(base + k * NS)] = v[b + B * k];
        }
        __syncthreads();
#pragma unroll
        for (unsigned m = 0; m < 16; ++m) v[m] = lds[padi(t + T * m)];
        __syncthreads();  // the next pass writes the same region
    }
}

template <int LOGN, unsigned p = 0>
__device__ __forceinline__ void fft_passes(float2 (&v)[16], unsigned t, const float2 *__restrict__ tw,
                                           float2 *lds) {
    using PL = FftPlan<LOGN>;
    if constexpr (p < PL::P) {
        constexpr bool last = p + 1 == PL::P;
        constexpr unsigned R = last ? PL::LAST : 16u;
        fft_pass<LOGN, R, (1u << (4 * p)), last>(v, t, tw, lds);
        fft_passes<LOGN, p + 1>(v, t, tw, lds);
    }
}

// Rows filtered with one real, even spectrum: two rows per complex
// transform, F transforms per workgroup.  FDK / Parker pre-weights are
// applied on load; the output is conj(FFT(conj(H . FFT(x + i y)))) / N,
// i.e. row a = Re, row a + 1 = -Im of FFT(conj(H . Z)) / N.
template <int LOGN, unsigned WGT>
__global__ __launch_bounds__((FftPlan<LOGN, WGT>::WG)) void filter_rows_stockham(FilterArgs A) {
    using PL = FftPlan<LOGN, WGT>;
    constexpr unsigned N = PL::N, T = PL::T;
    __shared__ float2 lds_all[PL::F * PL::PADN];
    const unsigned f = threadIdx.x / T, t = threadIdx.x % T;
    float2 *lds = lds_all + f * PL::PADN;
    const size_t r0 = ((size_t)blockIdx.x * PL::F + f) * 2;
    const bool has0 = r0 < A.n_rows, has1 = r0 + 1 < A.n_rows;
    const unsigned nd = A.n_det;
    float2 v[16];
    {
        // Every load is issued before any loaded value is used: the loads
        // sit in wave-uniform branches (column block m inside the row) with
        // clamped indices, and the products and the zero padding come after
        // all of them.  (Multiplying inside per-element branches made the
        // compiler wait for each weight before the next load.)  Rows past
        // the end read a valid row and are zeroed by the selects.
        const float *row0 = A.rows + (has0 ? r0 : 0) * nd;
        const float *row1 = has1 ? row0 + nd : row0;
        const bool wt = A.weights != nullptr, wt2 = A.weights2 != nullptr;
        const float *w0 = wt ? A.weights + ((has0 ? r0 : 0) / A.rows_per_weight) * nd : nullptr;
        const float *w1 = wt ? A.weights + ((has1 ? r0 + 1 : 0) / A.rows_per_weight) * nd : nullptr;
        const float *p0 = wt2 ? A.weights2 + ((has0 ? r0 : 0) % A.w2_rows) * nd : nullptr;
        const float *p1 = wt2 ? A.weights2 + ((has1 ? r0 + 1 : 0) % A.w2_rows) * nd : nullptr;
        float2 wv[16], pv[16];
#pragma unroll
        for (unsigned m = 0; m < 16; ++m) {
            v[m] = wv[m] = pv[m] = make_float2(0.f, 0.f);
            if (T * m < nd) {
                const unsigned i = min(t + T * m, nd - 1);
                v[m] = make_float2(row0[i], row1[i]);
                if (wt) wv[m] = make_float2(w0[i], w1[i]);
                if (wt2) pv[m] = make_float2(p0[i], p1[i]);
            }
        }
#pragma unroll
        for (unsigned m = 0; m < 16; ++m) {
            float a = v[m].x, b = v[m].y;
            if (wt) {
                a *= wv[m].x;
                b *= wv[m].y;
            }
            if (wt2) {
                a *= pv[m].x;
                b *= pv[m].y;
            }
            const bool in = t + T * m < nd;
            v[m] = make_float2(in && has0 ? a : 0.f, in && has1 ? b : 0.f);
        }
    }
    fft_passes<LOGN>(v, t, A.tw, lds);
    // x H, conjugated (inverse as conj(FFT(conj .)))
#pragma unroll
    for (unsigned m = 0; m < 16; ++m) {
        const float h = A.spec[t + T * m];
        v[m] = make_float2(v[m].x * h, -v[m].y * h);
    }
    fft_passes<LOGN>(v, t, A.tw, lds);
    const float inv = 1.0f / (float)N;
    float *row0 = A.rows + r0 * nd;
    float *row1 = row0 + nd;
#pragma unroll
    for (unsigned m = 0; m < 16; ++m) {
        const unsigned i = t + T * m;
        if (i < nd) {
            if (has0) row0[i] = v[m].x * inv;
            if (has1) row1[i] = -v[m].y * inv;
        }
    }
}

template <int LOGN, unsigned WGT>
void launch_stockham_wg(const FilterArgs &A, hipStream_t s) {
    using PL = FftPlan<LOGN, WGT>;
    const size_t per = 2 * (size_t)PL::F;  // rows per workgroup
    const size_t groups = (A.n_rows + per - 1) / per;
    AMX_REQUIRE(groups < (1ull << 31), "filter: too many rows");
    filter_rows_stockham<LOGN, WGT><<<(unsigned)groups, PL::WG, 0, s>>>(A);
    AMX_LAUNCHED();
}

// Workgroup size of the Stockham filter (threads, several transforms per
// group when one takes fewer): ASTRA_AMD_FILTER_WG = 64 (default), 128,
// 256 or 512, read at every call.  Measured at 768 x 720 rows x 2048
// points: 2.08 / 2.08 / 2.20 / 2.38 ms; at 1536 x 1440 x 4096: 17.2 /
// 17.3 / 17.3 / 19.7 ms.
template <int LOGN>
void launch_stockham(const FilterArgs &A, hipStream_t s) {
    const char *e = std::getenv("ASTRA_AMD_FILTER_WG");
    const int wg = e ? std::atoi(e) : 64;
    if (wg == 64) launch_stockham_wg<LOGN, 64>(A, s);
    else if (wg == 128) launch_stockham_wg<LOGN, 128>(A, s);
    else if (wg == 512) launch_stockham_wg<LOGN, 512>(A, s);
    else launch_stockham_wg<LOGN, 256>(A, s);
}

// transforms of 16 .. 4096 points (detectors up to 2048 bins; at most
// 34 KiB of static LDS per workgroup)
bool stockham_launch(const FilterArgs &A, unsigned log2n, hipStream_t s) {
    switch (log2n) {
    case 4: launch_stockham<4>(A, s); return true;
    case 5: launch_stockham<5>(A, s); return true;
    case 6: launch_stockham<6>(A, s); return true;
    case 7: launch_stockham<7>(A, s); return true;
    case 8: launch_stockham<8>(A, s); return true;
    case 9: launch_stockham<9>(A, s); return true;
    case 10: launch_stockham<10>(A, s); return true;
    case 11: launch_stockham<11>(A, s); return true;
    case 12: launch_stockham<12>(A, s); return true;
    default: return false;  // 8192 points and up: the LDS kernels
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
        float a = 0.f, b = 0.f, wa = 1.f, wb = 1.f, va = 1.f, vb = 1.f;
        const unsigned j = 2 * m;
        if (j < A.n_det) {
            a = row[j];
            if (w) wa = w[j];
            if (v) va = v[j];
        }
        if (j + 1 < A.n_det) {
            b = row[j + 1];
            if (w) wb = w[j + 1];
            if (v) vb = v[j + 1];
        }
        // loads first, then the (exact where a weight is absent) products
        Z[__brev(m) >> shift] = make_float2(a * wa * va, b * wb * vb);
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
        // loads first, then the (exact where a weight is absent) products
        const float v = row[i], wa = a ? a[i] : 1.f, wb = b ? b[i] : 1.f;
        row[i] = v * wa * wb;
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
    // ASTRA_AMD_FILTER_LDS_RADIX4=1 keeps the LDS radix-2^2 kernel for the
    // real-spectrum case (read at every call; for comparing the two)
    const char *old_path = std::getenv("ASTRA_AMD_FILTER_LDS_RADIX4");
    if (!f.general && !(old_path && old_path[0] == '1')) {
        ProfToken pt = prof_begin("filter_rows", (double)n_rows, s);
        const bool done = stockham_launch(A, l, s);
        prof_end(pt, s);
        if (done) return;
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
