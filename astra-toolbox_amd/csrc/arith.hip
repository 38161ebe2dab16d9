// Elementwise and reduction kernels used by the algorithm objects: the
// reference's processData<op> functors (cuda/2d/arith.cu:43-158) that the
// SIRT/CGLS loops need, and float64-accumulated dot products
// (dotProduct2D accumulates in fp32, cuda/2d/util.cu:231-357).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "devutil.hpp"
#include "internal.hpp"

namespace amx {

namespace {
inline unsigned grid_for(size_t n) {
    size_t g = (n + 255) / 256;
    return (unsigned)std::min<size_t>(std::max<size_t>(g, 1), 65536);
}
}  // namespace

__global__ void k_fill(float *p, size_t n, float v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// opInvert (cuda/2d/arith.cu:99-106) followed by a scale
__global__ void k_invert(float *p, size_t n, float scale) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        float x = p[i];
        p[i] = (x > 0.000001f) ? scale / x : 0.f;
    }
}

__global__ void k_mul(float *p, const float *q, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] *= q[i];
}

__global__ void k_copy(float *d, const float *s, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        d[i] = s[i];
}

// opClampMin / opClampMax / opClampMinMask / opClampMaxMask
__global__ void k_clamp(float *p, size_t n, float lo, float hi, const float *lom, const float *him) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        float x = fminf(fmaxf(p[i], lo), hi);
        if (lom) x = fmaxf(x, lom[i]);
        if (him) x = fminf(x, him[i]);
        p[i] = x;
    }
}

// rows[r][i] *= w[r / rpw][i]; one workgroup per row
__global__ void k_rowscale(float *rows, unsigned len, const float *w, size_t rpw) {
    float *row = rows + (size_t)blockIdx.x * len;
    const float *wr = w + (size_t)(blockIdx.x / rpw) * len;
    for (unsigned i = threadIdx.x; i < len; i += blockDim.x) row[i] *= wr[i];
}

// block partial sums of a[i]*b[i] in float64
__global__ __launch_bounds__(256) void k_dot_partial(const float *a, const float *b, size_t n,
                                                     double *part) {
    __shared__ double sm[256];
    double acc = 0.0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        acc += (double)a[i] * (double)b[i];
    sm[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sm[threadIdx.x] += sm[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

__global__ __launch_bounds__(256) void k_dot_final(const double *part, unsigned np, double *out) {
    __shared__ double sm[256];
    double acc = 0.0;
    for (unsigned i = threadIdx.x; i < np; i += blockDim.x) acc += part[i];
    sm[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sm[threadIdx.x] += sm[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sm[0];
}

// y += sign * (num / den) * x   (opAddScaled with a device-side scalar)
__global__ void k_axpy_ratio(float *y, const float *x, size_t n, const double *num,
                             const double *den, double sign) {
    const double d = *den;
    const float alpha = (float)(d != 0.0 ? sign * (*num / d) : 0.0);
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        y[i] = fmaf(alpha, x[i], y[i]);
}

// p = z + (num / den) * p   (opScaleAndAdd with a device-side scalar)
__global__ void k_xpby_ratio(float *p, const float *z, size_t n, const double *num,
                             const double *den) {
    const double d = *den;
    const float beta = (float)(d != 0.0 ? (*num / d) : 0.0);
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = fmaf(beta, p[i], z[i]);
}

void fill_launch(float *p, size_t n, float v, hipStream_t s) {
    if (!n) return;
    k_fill<<<grid_for(n), 256, 0, s>>>(p, n, v);
    AMX_LAUNCHED();
}
void invert_launch(float *p, size_t n, float scale, hipStream_t s) {
    if (!n) return;
    k_invert<<<grid_for(n), 256, 0, s>>>(p, n, scale);
    AMX_LAUNCHED();
}
void mul_launch(float *p, const float *q, size_t n, hipStream_t s) {
    if (!n) return;
    k_mul<<<grid_for(n), 256, 0, s>>>(p, q, n);
    AMX_LAUNCHED();
}
void copy_launch(float *d, const float *src, size_t n, hipStream_t s) {
    if (!n) return;
    k_copy<<<grid_for(n), 256, 0, s>>>(d, src, n);
    AMX_LAUNCHED();
}
void clamp_launch(float *p, size_t n, float lo, float hi, const float *lom, const float *him,
                  hipStream_t s) {
    if (!n) return;
    k_clamp<<<grid_for(n), 256, 0, s>>>(p, n, lo, hi, lom, him);
    AMX_LAUNCHED();
}

void rowscale_launch(float *rows, size_t n_rows, unsigned len, const float *w, size_t rpw,
                     hipStream_t s) {
    if (!n_rows) return;
    AMX_REQUIRE(n_rows < (1ull << 31), "rowscale: too many rows");
    k_rowscale<<<(unsigned)n_rows, 256, 0, s>>>(rows, len, w, rpw ? rpw : 1);
    AMX_LAUNCHED();
}

// out (device double) = sum a*b; uses a stream-ordered partials buffer
void dot_launch(const float *a, const float *b, size_t n, double *out, hipStream_t s) {
    const unsigned g = std::min<unsigned>(grid_for(n), 1024);
    Scratch part(sizeof(double) * g, s);
    k_dot_partial<<<g, 256, 0, s>>>(a, b, n, part.as<double>());
    AMX_LAUNCHED();
    k_dot_final<<<1, 256, 0, s>>>(part.as<double>(), g, out);
    AMX_LAUNCHED();
}

void axpy_ratio_launch(float *y, const float *x, size_t n, const double *num, const double *den,
                       double sign, hipStream_t s) {
    if (!n) return;
    k_axpy_ratio<<<grid_for(n), 256, 0, s>>>(y, x, n, num, den, sign);
    AMX_LAUNCHED();
}

void xpby_ratio_launch(float *p, const float *z, size_t n, const double *num, const double *den,
                       hipStream_t s) {
    if (!n) return;
    k_xpby_ratio<<<grid_for(n), 256, 0, s>>>(p, z, n, num, den);
    AMX_LAUNCHED();
}

}  // namespace amx
