"""Oracle SIRT / CGLS in float64 over oracle projectors.

Test infrastructure only (see ``oracle/__init__.py``).  ``fp(x)`` and
``bp(y)`` are closures over the oracle projectors with the geometry's output
scale already applied.
"""
import numpy as np

EPS = 1e-6


def _inv(x):
    """``opInvert`` (cuda/2d/arith.cu:99-106): 1/x if x > 1e-6 else 0."""
    out = np.zeros_like(x)
    m = x > EPS
    out[m] = 1.0 / x[m]
    return out


def sirt(fp, bp, sino, x0, iterations, relaxation=1.0, min_c=None, max_c=None):
    """GPU SIRT (src/CudaSirtAlgorithm.cpp:187-222, 347-379;
    src/CudaSirtAlgorithm3D.cpp:261-376): lineW = inv(FP(1)),
    pixelW = lambda inv(BP(1)); r = (b - Ax) lineW; x += pixelW BP(r);
    clamp."""
    x = np.array(x0, dtype=np.float64)
    b = np.asarray(sino, dtype=np.float64)
    line_w = _inv(fp(np.ones_like(x)))
    pixel_w = relaxation * _inv(bp(np.ones_like(b)))
    for _ in range(iterations):
        r = (b - fp(x)) * line_w
        x = x + pixel_w * bp(r)
        if min_c is not None:
            x = np.maximum(x, min_c)
        if max_c is not None:
            x = np.minimum(x, max_c)
    return x


def cgls(fp, bp, sino, x0, iterations):
    """GPU CGLS (src/CudaCglsAlgorithm.cpp:207-266): restarts on every run."""
    x = np.array(x0, dtype=np.float64)
    b = np.asarray(sino, dtype=np.float64)
    r = b - fp(x)
    p = bp(r)
    gamma = float(np.sum(p * p))
    for _ in range(iterations):
        w = fp(p)
        alpha = gamma / float(np.sum(w * w))
        x = x + alpha * p
        r = r - alpha * w
        z = bp(r)
        gamma_new = float(np.sum(z * z))
        beta = gamma_new / gamma
        gamma = gamma_new
        p = z + beta * p
    return x


def residual_norm(fp, sino, x):
    """``getResidualNorm``: ||b - Ax|| (src/CudaSirtAlgorithm.cpp:399-425)."""
    r = np.asarray(sino, dtype=np.float64) - fp(x)
    return float(np.sqrt(np.sum(r * r)))


def em(fp, bp, sino, x0, iterations):
    """GPU EM (src/CudaEMAlgorithm.cpp:149-253): pixelW = inv(BP(1));
    x *= pixelW BP(b / FP(x)) with the ratio 0 where FP(x) <= 1e-6
    (cuda/2d/arith.cu:91-98)."""
    x = np.array(x0, dtype=np.float64)
    b = np.asarray(sino, dtype=np.float64)
    pixel_w = _inv(bp(np.ones_like(b)))
    for _ in range(iterations):
        t = fp(x)
        ratio = np.where(t > 1e-6, b / np.where(t > 1e-6, t, 1.0), 0.0)
        x = x * (bp(ratio) * pixel_w)
    return x


def sart(fp_angle, bp_angle, sino, x0, order, relaxation=1.0, min_c=None, max_c=None,
         line_w=None):
    """GPU SART (src/CudaSartAlgorithm.cpp:196-316): for each angle a of the
    order, r = (b_a - A_a x) lineW_a; x += lambda A_a^T r (SART's unweighted
    BP); clamp.  fp_angle(x, a) -> row a of the FP; bp_angle(row, a) ->
    volume; line_w: inv(FP(1)) rows."""
    x = np.array(x0, dtype=np.float64)
    b = np.asarray(sino, dtype=np.float64)
    for a in order:
        r = (b[a] - fp_angle(x, a)) * line_w[a]
        x = x + relaxation * bp_angle(r, a)
        if min_c is not None:
            x = np.maximum(x, min_c)
        if max_c is not None:
            x = np.minimum(x, max_c)
    return x
