/*
 * C restatement of the reference CPU path used for the CPU baseline:
 * the parallel-beam 'linear' projector and CPU SIRT.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/ and bench.py's cpu_baseline leg
 * as the checker / baseline; the product never links it.
 *
 * Follows
 *   genParProjections               src/GeometryUtil2D.cpp:40-79
 *   projectBlock_internal           include/astra/ParallelBeamLinearKernelProjector2D.inl:128-237
 *   DefaultFP/BP, DiffFP, TotalPixelWeight, TotalRayLength policies
 *                                   include/astra/DataProjectorPolicies.inl:36-176, 299-374
 *   CSirtAlgorithm::run             src/SirtAlgorithm.cpp:161-286
 * in float32 with the same loop order, so results match the reference CPU
 * build up to the compiler's floating-point contraction choices.  The
 * projector body uses the float overloads of sqrt/abs/floor, as the
 * reference (which includes the .inl after `using namespace std`).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    float rx, ry, dsx, dsy, ux, uy;
} ParVec;

typedef struct {
    int na, nd, rows, cols;
    float minx, maxx, miny, maxy;
    ParVec *v;
} Geo;

enum { MODE_FP = 0, MODE_DIFF = 1, MODE_DIFF_WEIGHTS = 2, MODE_BP = 3 };

/* genParProjections: base vectors rotated in double, stored as float */
static void gen_par(int na, int nd, float w, const float *angles, ParVec *out) {
    const float bdsx = (float)(nd * (double)w * -0.5f);
    for (int i = 0; i < na; ++i) {
        double a = angles[i], c = cos(a), s = sin(a);
        out[i].rx = (float)(0.0f * c - 1.0f * s);
        out[i].ry = (float)(0.0f * s + 1.0f * c);
        out[i].dsx = (float)(bdsx * c - 0.0f * s);
        out[i].dsy = (float)(bdsx * s + 0.0f * c);
        out[i].ux = (float)(w * c - 0.0f * s);
        out[i].uy = (float)(w * s + 0.0f * c);
    }
}

/* One pass of the linear-kernel projector with the policy selected by mode:
 *   FP:   sino[r] = 0; sino[r] += vol[p] w
 *   DIFF: diff[r] = base[r]; diff[r] -= vol[p] w  (+ pixw[p] += w, rayl[r] += w)
 *   BP:   vol[p] += sino[r] w                                              */
static void project(const Geo *g, int mode, float *sino, const float *base, float *vol,
                    float *pixw, float *rayl) {
    const float px = (g->maxx - g->minx) / (float)g->cols;
    const float py = (g->maxy - g->miny) / (float)g->rows;
    const float ipx = 1.0f / px, ipy = 1.0f / py;
    const int cols = g->cols, rows = g->rows;
    for (int a = 0; a < g->na; ++a) {
        const ParVec *p = &g->v[a];
        const int vertical = fabsf(p->rx) < fabsf(p->ry);
        const float Ex = g->minx + px * 0.5f;
        const float Ey = g->maxy - py * 0.5f;
        for (int d = 0; d < g->nd; ++d) {
            const int ray = a * g->nd + d;
            if (mode == MODE_FP) sino[ray] = 0.0f;
            else if (mode == MODE_DIFF || mode == MODE_DIFF_WEIGHTS) sino[ray] = base[ray];
            const float Dx = p->dsx + (d + 0.5f) * p->ux;
            const float Dy = p->dsy + (d + 0.5f) * p->uy;
            int isin = 0;
#define WEIGHT(idx, wt)                                                     \
    do {                                                                    \
        float w_ = (wt);                                                    \
        if (mode == MODE_FP) sino[ray] += vol[idx] * w_;                    \
        else if (mode == MODE_BP) vol[idx] += sino[ray] * w_;               \
        else {                                                              \
            sino[ray] -= vol[idx] * w_;                                     \
            if (mode == MODE_DIFF_WEIGHTS) { pixw[idx] += w_; rayl[ray] += w_; } \
        }                                                                   \
    } while (0)
            if (vertical) {
                const float RxOverRy = p->rx / p->ry;
                const float lpr = px * sqrtf(p->ry * p->ry + p->rx * p->rx) / fabsf(p->ry);
                const float deltac = -py * RxOverRy * ipx;
                float c = (Dx + (Ey - Dy) * RxOverRy - Ex) * ipx;
                for (int row = 0; row < rows; ++row, c += deltac) {
                    int col = (int)floorf(c);
                    if (col < -1 || col >= cols) {
                        if (!isin) continue;
                        break;
                    }
                    float off = c - (float)col;
                    int idx = row * cols + col;
                    if (col >= 0) WEIGHT(idx, (1.0f - off) * lpr);
                    idx++;
                    if (col + 1 < cols) WEIGHT(idx, off * lpr);
                    isin = 1;
                }
            } else {
                const float RyOverRx = p->ry / p->rx;
                const float lpc = py * sqrtf(p->ry * p->ry + p->rx * p->rx) / fabsf(p->rx);
                const float deltar = -px * RyOverRx * ipy;
                float r = -(Dy + (Ex - Dx) * RyOverRx - Ey) * ipy;
                for (int col = 0; col < cols; ++col, r += deltar) {
                    int row = (int)floorf(r);
                    if (row < -1 || row >= rows) {
                        if (!isin) continue;
                        break;
                    }
                    float off = r - (float)row;
                    int idx = row * cols + col;
                    if (row >= 0) WEIGHT(idx, (1.0f - off) * lpc);
                    idx += cols;
                    if (row + 1 < rows) WEIGHT(idx, off * lpc);
                    isin = 1;
                }
            }
#undef WEIGHT
        }
    }
}

static int make_geo(Geo *g, int rows, int cols, int na, int nd, float w, const float *angles) {
    g->na = na;
    g->nd = nd;
    g->rows = rows;
    g->cols = cols;
    g->minx = -cols / 2.0f;
    g->maxx = cols / 2.0f;
    g->miny = -rows / 2.0f;
    g->maxy = rows / 2.0f;
    g->v = (ParVec *)malloc(sizeof(ParVec) * (size_t)na);
    if (!g->v) return -1;
    gen_par(na, nd, w, angles, g->v);
    return 0;
}

/* CForwardProjectionAlgorithm::run: sino = A vol */
int cpu_linear_fp(int rows, int cols, int na, int nd, float w, const float *angles,
                  const float *vol, float *sino) {
    Geo g;
    if (make_geo(&g, rows, cols, na, nd, w, angles)) return -1;
    project(&g, MODE_FP, sino, NULL, (float *)vol, NULL, NULL);
    free(g.v);
    return 0;
}

/* CBackProjectionAlgorithm::run: vol = A^T sino */
int cpu_linear_bp(int rows, int cols, int na, int nd, float w, const float *angles,
                  const float *sino, float *vol) {
    Geo g;
    if (make_geo(&g, rows, cols, na, nd, w, angles)) return -1;
    memset(vol, 0, sizeof(float) * (size_t)rows * cols);
    project(&g, MODE_BP, (float *)sino, NULL, vol, NULL, NULL);
    free(g.v);
    return 0;
}

/* CSirtAlgorithm::run(iters) on a fresh algorithm object; vol is the
 * initial iterate and receives the result. */
int cpu_linear_sirt(int rows, int cols, int na, int nd, float w, const float *angles,
                    const float *sino, float *vol, int iters, float lambda) {
    const float eps = 1e-6f; /* include/astra/Globals.h:112 */
    Geo g;
    if (make_geo(&g, rows, cols, na, nd, w, angles)) return -1;
    const size_t nv = (size_t)rows * cols, ns = (size_t)na * nd;
    float *diff = (float *)malloc(sizeof(float) * ns);
    float *rayl = (float *)calloc(ns, sizeof(float));
    float *pixw = (float *)calloc(nv, sizeof(float));
    float *tmp = (float *)malloc(sizeof(float) * nv);
    if (!diff || !rayl || !pixw || !tmp) return -1;
    if (iters > 0) {
        project(&g, MODE_DIFF_WEIGHTS, diff, sino, vol, pixw, rayl);
        for (size_t i = 0; i < nv; ++i) {
            float x = pixw[i];
            x = (x < -eps || x > eps) ? 1.0f / x : 0.0f;
            pixw[i] = lambda * x;
        }
        for (size_t i = 0; i < ns; ++i) {
            float x = rayl[i];
            rayl[i] = (x < -eps || x > eps) ? 1.0f / x : 0.0f;
        }
    }
    for (int it = 0; it < iters; ++it) {
        if (it > 0) project(&g, MODE_DIFF, diff, sino, vol, NULL, NULL);
        for (size_t i = 0; i < ns; ++i) diff[i] *= rayl[i];
        memset(tmp, 0, sizeof(float) * nv);
        project(&g, MODE_BP, diff, NULL, tmp, NULL, NULL);
        for (size_t i = 0; i < nv; ++i) tmp[i] *= pixw[i];
        for (size_t i = 0; i < nv; ++i) vol[i] += tmp[i];
    }
    free(diff);
    free(rayl);
    free(pixw);
    free(tmp);
    free(g.v);
    return 0;
}
