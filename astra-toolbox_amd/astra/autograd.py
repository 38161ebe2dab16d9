"""Differentiable projectors and unrolled SIRT for PyTorch.

No counterpart in the reference (its projectors are not differentiable and
its GPU back projectors are not exact adjoints).  Here every forward
projector has an exactly matched adjoint: the back projection used for
gradients applies the transpose of the forward projector's own sampling
(the same rays, sample positions and bilinear / linear weights,
scattered instead of gathered), so that <A x, y> = <x, A^T y> up to fp32
rounding.

* ``forward_project(vol, projector)``: A vol; gradient A^T g (matched).
* ``back_project(proj, projector)``: A^T proj (matched); gradient A g.
* ``sirt_unrolled(b, projector, iterations, ...)``: x_N of SIRT built on the
  matched pair, differentiable in b and x0.  The backward pass recomputes
  the iterates from checkpoints (every ~sqrt(N) iterations) instead of
  storing all N, and runs the adjoint recursion with the same kernels.

``projector`` is a projector id from ``astra.create_projector`` (2D
'cuda' or 3D 'cuda3d') or a projector object.  Tensors are float32 on the
projector's GPU; shapes are the astra layouts (2D volume [rows][cols],
sinogram [angles][dets]; 3D volume [z][y][x], projections [v][angles][u]).
"""
import math

import torch

from . import _core, _lib, projector as _p2, projector3d as _p3
from .log import AstraError


def _resolve(projector):
    if isinstance(projector, _core.Projector):
        return projector
    try:
        return _p3.get(projector)
    except Exception:
        return _p2.get(projector)


def _check(t, shape, what):
    if not isinstance(t, torch.Tensor) or t.device.type != "cuda":
        raise AstraError("%s must be a GPU tensor" % what)
    if t.dtype != torch.float32:
        raise AstraError("%s must be float32" % what)
    if tuple(t.shape) != tuple(shape):
        raise AstraError("%s has shape %s, expected %s" % (what, tuple(t.shape), tuple(shape)))
    return t.contiguous()


def _fp(P, vol, out=None):
    if out is None:
        out = torch.empty(P.proj_shape, dtype=torch.float32, device=vol.device)
    _core.fp(P, vol, out, _lib.MODE_SET)
    return out


def _bp_matched(P, proj):
    out = torch.empty(P.vol_shape, dtype=torch.float32, device=proj.device)
    _core.bp(P, proj, out, _lib.MODE_SET, kernel=_lib.KER3D_MATCHED_BP)
    return out


class _ForwardProject(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vol, P):
        ctx.P = P
        return _fp(P, vol)

    @staticmethod
    def backward(ctx, g):
        return _bp_matched(ctx.P, _check(g, ctx.P.proj_shape, "gradient")), None


class _BackProject(torch.autograd.Function):
    @staticmethod
    def forward(ctx, proj, P):
        ctx.P = P
        return _bp_matched(P, proj)

    @staticmethod
    def backward(ctx, g):
        return _fp(ctx.P, _check(g, ctx.P.vol_shape, "gradient")), None


def forward_project(vol, projector):
    """A vol (the projector's forward projection), differentiable."""
    P = _resolve(projector)
    return _ForwardProject.apply(_check(vol, P.vol_shape, "volume"), P)


def back_project(proj, projector):
    """A^T proj with the exactly matched adjoint, differentiable."""
    P = _resolve(projector)
    return _BackProject.apply(_check(proj, P.proj_shape, "projection data"), P)


def _inv(t):
    # 1/x where x > 1e-6, else 0 (cuda/2d/arith.cu:99-106)
    return torch.where(t > 1e-6, 1.0 / t.clamp_min(1e-30), torch.zeros_like(t))


def sirt_weights(projector, relaxation=1.0, device=None):
    """(R, C) of SIRT on the matched pair: R = inv(A 1), C = relaxation *
    inv(A^T 1) (src/CudaSirtAlgorithm.cpp:187-222)."""
    P = _resolve(projector)
    dev = device or P.device()
    R = _inv(_fp(P, torch.ones(P.vol_shape, dtype=torch.float32, device=dev)))
    C = _inv(_bp_matched(P, torch.ones(P.proj_shape, dtype=torch.float32, device=dev))) * relaxation
    return R, C


def _constrained(lo, hi):
    return lo != -math.inf or hi != math.inf


def _step(P, R, C, b, x, lo, hi, pbuf=None):
    """One SIRT iteration; returns (x_next, z) with z the unclamped value
    (the same tensor when there are no constraints).  pbuf: optional
    projection-shaped work buffer for the residual (overwritten)."""
    r = _fp(P, x, pbuf)
    torch.sub(b, r, out=r)      # b - A x
    r.mul_(R)
    z = _bp_matched(P, r)
    z.mul_(C).add_(x)           # x + C A^T R (b - A x)
    return (z.clamp(lo, hi) if _constrained(lo, hi) else z), z


def _step_z(P, R, C, b, x, pbuf):
    """z = x + C A^T R (b - A x), the unclamped SIRT update (a new volume)."""
    r = _fp(P, x, pbuf)
    torch.sub(b, r, out=r)
    r.mul_(R)
    z = _bp_matched(P, r)
    return z.mul_(C).add_(x)


class _UnrolledSIRT(torch.autograd.Function):
    @staticmethod
    def forward(ctx, b, x0, P, iterations, R, C, lo, hi, every, keep_segment):
        x = x0
        # without constraints the backward pass needs no iterate: keep none
        cons = _constrained(lo, hi)
        ckpts = [x0] if cons else []
        pbuf = torch.empty_like(b)  # one residual buffer for every iteration
        for k in range(iterations):
            if cons and k > 0 and k % every == 0:
                ckpts.append(x)
            x, _ = _step(P, R, C, b, x, lo, hi, pbuf)
        del pbuf
        ctx.P, ctx.n, ctx.lo, ctx.hi, ctx.every = P, iterations, lo, hi, every
        ctx.keep_segment = keep_segment
        ctx.save_for_backward(b, R, C, *ckpts)
        return x

    @staticmethod
    def backward(ctx, g):
        b, R, C, *ckpts = ctx.saved_tensors
        P, n, lo, hi, every = ctx.P, ctx.n, ctx.lo, ctx.hi, ctx.every
        g = _check(g, P.vol_shape, "gradient").clone()
        gb = torch.zeros_like(b)
        pbuf = torch.empty_like(b)  # the only projection-shaped work buffer
        cons = _constrained(lo, hi)
        cgbuf = None if cons else torch.empty_like(g)  # C h (constrained: in z's memory)

        def iterate_from(x, steps):
            # x_{j+1} = clamp(z_j), clamped in place: two volumes at a time
            for _ in range(steps):
                x = _step_z(P, R, C, b, x, pbuf)
                if cons:
                    x.clamp_(lo, hi)
            return x

        nseg = len(ckpts) if cons else 1
        for seg in range(nseg - 1, -1, -1):
            start, stop = (seg * every, min(n, (seg + 1) * every)) if cons else (0, n)
            # keep_segment: recompute x_start .. x_{stop-1} once and hold
            # them; else recompute each x_k from the checkpoint when needed
            # (one volume instead of `every`, for about every / 2 times the
            # recomputation)
            xs = [ckpts[seg]] if cons else []
            if cons and ctx.keep_segment:
                for _ in range(start, stop - 1):
                    xs.append(_step(P, R, C, b, xs[-1], lo, hi, pbuf)[0])
            for k in range(stop - 1, start - 1, -1):
                # x_{k+1} = clamp(z), z = x_k + C A^T R (b - A x_k)
                if cons:  # torch.clamp's gradient: h = g where lo <= z <= hi
                    xk = xs[k - start] if ctx.keep_segment else iterate_from(ckpts[seg], k - start)
                    z = _step_z(P, R, C, b, xk, pbuf)
                    del xk
                    g.masked_fill_(~((z >= lo) & (z <= hi)), 0.0)
                    cg = torch.mul(C, g, out=z)  # z's memory is free again
                    del z
                else:
                    cg = torch.mul(C, g, out=cgbuf)
                t = _fp(P, cg, pbuf)           # dz/db^T h = R A C h
                t.mul_(R)
                gb += t
                del cg
                g.sub_(_bp_matched(P, t))      # dz/dx^T h = h - A^T R A C h
                if cons and ctx.keep_segment:
                    xs[k - start] = None       # x_k is no longer needed
            del xs
        return gb, g, None, None, None, None, None, None, None, None


def sirt_unrolled(b, projector, iterations, x0=None, relaxation=1.0, min_constraint=None,
                  max_constraint=None, checkpoint_every=None, weights=None, keep_segment=False):
    """x_N of `iterations` SIRT steps x <- clamp(x + C A^T R (b - A x)) on the
    matched projector pair, differentiable in b and x0.  The backward pass
    keeps only every `checkpoint_every`-th iterate and recomputes the
    others: each from its checkpoint when it is needed (default, with
    three checkpoints; the least memory), or once per segment, holding the
    segment's iterates (keep_segment=True, ~sqrt(N) checkpoints; less
    recomputation).  Only the
    constrained iteration needs the iterates in the backward pass (for the
    clamp masks): without constraints the gradient is the same linear
    recursion for every k and nothing is recomputed."""
    P = _resolve(projector)
    b = _check(b, P.proj_shape, "projection data")
    if x0 is None:
        x0 = torch.zeros(P.vol_shape, dtype=torch.float32, device=b.device)
    x0 = _check(x0, P.vol_shape, "initial volume")
    iterations = int(iterations)
    if iterations < 1:
        raise AstraError("sirt_unrolled: iterations must be >= 1")
    # default checkpoints: ~sqrt(N) segments when each is held at once,
    # else three (x_0 and two more), the least memory for a bounded
    # recomputation (about N^2 / 18 extra steps)
    if checkpoint_every:
        every = int(checkpoint_every)
    elif keep_segment:
        every = max(1, math.ceil(math.sqrt(iterations)))
    else:
        every = max(1, math.ceil(iterations / 3))
    R, C = weights if weights is not None else sirt_weights(P, relaxation, b.device)
    lo = -math.inf if min_constraint is None else float(min_constraint)
    hi = math.inf if max_constraint is None else float(max_constraint)
    return _UnrolledSIRT.apply(b, x0, P, iterations, R, C, lo, hi, every, bool(keep_segment))


def sirt_unrolled_stored(b, projector, iterations, x0=None, relaxation=1.0, min_constraint=None,
                         max_constraint=None, weights=None):
    """The same iteration written with differentiable forward_project /
    back_project, so autograd stores every iterate (reference for the
    recomputing version, and the memory it saves)."""
    P = _resolve(projector)
    if x0 is None:
        x0 = torch.zeros(P.vol_shape, dtype=torch.float32, device=b.device)
    R, C = weights if weights is not None else sirt_weights(P, relaxation, b.device)
    lo = -math.inf if min_constraint is None else float(min_constraint)
    hi = math.inf if max_constraint is None else float(max_constraint)
    x = x0
    for _ in range(int(iterations)):
        x = torch.clamp(x + C * back_project((b - forward_project(x, P)) * R, P), lo, hi)
    return x
