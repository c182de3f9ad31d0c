"""Adjoint entry points and their torch.autograd bindings.

The adjoint formulas themselves are pinned on the CPU against explicit
transposed matrices (test_oracle.py); here the GPU adjoints are compared
with them in FP64, held to the dot-product identity <A x, y> = <x, A^T y>
in FP32, and the autograd chain of the loss 1/2 |div(grad(invlap u)) - g|^2
is checked by central finite differences and by its closed form
(div grad invlap = P, the projection removing the modes every per-axis
derivative zeroes, so dL/du = P(P u - g) = P u - P g).
"""
import numpy as np
import pytest

from accfft_amd import fields
from conftest import rel_l2, rng

pytestmark = pytest.mark.gpu

SHAPES = [(16, 12, 10), (9, 10, 7), (64, 64, 64), (32, 8, 129)]


def _t(a, dev, dtype=None):
    import torch
    return torch.tensor(a, device=dev, dtype=dtype)


@pytest.mark.parametrize("n", SHAPES)
def test_adjoint_r2c_c2r_fp64(dev, orc, n):
    import torch
    import accfft_amd as af
    print(f"adjoint r2c/c2r FP64 {n}", flush=True)
    g = rng(50)
    hs = (n[0], n[1], n[2] // 2 + 1)
    G = g.standard_normal(hs) + 1j * g.standard_normal(hs)
    y = g.standard_normal(n)
    Gd, yd = _t(G, dev), _t(y, dev)
    x = torch.empty(n, dtype=torch.float64, device=dev)
    H = torch.empty(hs, dtype=torch.complex128, device=dev)
    plan = af.Plan("r2c", n, x, H, af.F64)
    G0, y0 = Gd.clone(), yd.clone()
    plan.adjoint_r2c(Gd, x)
    plan.adjoint_c2r(yd, H)
    torch.cuda.synchronize()
    assert torch.equal(Gd, G0) and torch.equal(yd, y0), "adjoint overwrote its input"
    assert rel_l2(x.cpu().numpy(), orc.adjoint_r2c(G, n)) < 1e-12
    assert rel_l2(H.cpu().numpy(), orc.adjoint_c2r(y)) < 1e-12
    plan.destroy()


@pytest.mark.parametrize("n", SHAPES)
def test_adjoint_grad_div_fp64(dev, orc, n):
    import torch
    import accfft_amd as af
    print(f"adjoint grad/div FP64 {n}", flush=True)
    g = rng(51)
    gs = [g.standard_normal(n) for _ in range(3)]
    v = g.standard_normal(n)
    gd = [_t(a, dev) for a in gs]
    vd = _t(v, dev)
    out = torch.empty(n, dtype=torch.float64, device=dev)
    plan = af.Plan("r2c", n, out, torch.empty(1, device=dev), af.F64)
    plan.adjoint_grad(*gd, out)
    o3 = [torch.empty_like(out) for _ in range(3)]
    plan.adjoint_divergence(vd, *o3)
    torch.cuda.synchronize()
    assert rel_l2(out.cpu().numpy(), orc.adjoint_grad(*gs)) < 1e-12
    for a, b in zip(o3, orc.adjoint_divergence(v)):
        assert rel_l2(a.cpu().numpy(), b) < 1e-12
    # a missing component contributes nothing; a missing output is skipped
    plan.adjoint_grad(gd[0], None, gd[2], out)
    oy = torch.full_like(out, 7.0)
    plan.adjoint_divergence(vd, None, oy, None)
    torch.cuda.synchronize()
    z = np.zeros(n)
    assert rel_l2(out.cpu().numpy(), orc.adjoint_grad(gs[0], z, gs[2])) < 1e-12
    assert rel_l2(oy.cpu().numpy(), orc.adjoint_divergence(v)[1]) < 1e-12
    plan.destroy()


def _dot(a, b):
    import torch
    if a.is_complex():
        return float((a.real.double() * b.real.double()).sum() + (a.imag.double() * b.imag.double()).sum())
    return float((a.double() * b.double()).sum())


@pytest.mark.parametrize("n", [(128, 128, 128), (64, 96, 130)])
def test_dot_product_identities_fp32(dev, n):
    """<A x, y> = <x, A^T y> for every map, FP32 plans (rel. 1e-5)."""
    import torch
    from accfft_amd.autograd import Spectral
    print(f"dot-product identities FP32 {n}", flush=True)
    S = Spectral(n, torch.float32, dev)
    gen = torch.Generator(device=dev).manual_seed(52)

    def rr():
        return torch.randn(n, generator=gen, device=dev)

    def rc():
        return torch.randn(S.spectrum_shape, generator=gen, device=dev, dtype=torch.complex64)

    def check(lhs, rhs, scale, what):
        err = abs(lhs - rhs) / scale
        print(f"  {what}: rel {err:.2e}")
        assert err < 1e-5, what

    x, G = rr(), rc()
    Ax, ATG = S.r2c(x), S.r2c_adjoint(G)
    check(_dot(Ax, G), _dot(x, ATG), Ax.norm().item() * G.norm().item(), "r2c")
    H, y = rc(), rr()
    CH, CTy = S.c2r(H), S.c2r_adjoint(y)
    check(_dot(CH, y), _dot(H, CTy), CH.norm().item() * y.norm().item(), "c2r")
    u, gs = rr(), [rr() for _ in range(3)]
    gu = S.grad(u)
    lhs = sum(_dot(a, b) for a, b in zip(gu, gs))
    check(lhs, _dot(u, S.grad_adjoint(*gs)),
          sum(a.norm().item() * b.norm().item() for a, b in zip(gu, gs)), "grad")
    v = rr()
    dv = S.divergence(*gs)
    check(_dot(dv, v), sum(_dot(a, b) for a, b in zip(gs, S.divergence_adjoint(v))),
          dv.norm().item() * v.norm().item(), "divergence")
    for op in ("laplace", "inv_laplace", "biharmonic", "inv_biharmonic"):
        a, b = rr(), rr()
        La, Lb = S.op(op, a), S.op(op, b)
        check(_dot(La, b), _dot(a, Lb), La.norm().item() * b.norm().item(), op)


def test_gradcheck_fp64(dev):
    """torch.autograd.gradcheck (numerical Jacobians) on a small FP64 grid."""
    import torch
    from accfft_amd import autograd as ag
    n = (6, 4, 5)
    S = ag.Spectral(n, torch.float64, dev)
    gen = torch.Generator(device=dev).manual_seed(53)
    x = torch.randn(n, generator=gen, device=dev, dtype=torch.float64, requires_grad=True)
    G = torch.randn(S.spectrum_shape, generator=gen, device=dev, dtype=torch.complex128,
                    requires_grad=True)
    kw = dict(eps=1e-6, atol=1e-8, rtol=1e-6)
    assert torch.autograd.gradcheck(lambda t: ag.rfft3(t, S), (x,), **kw)
    assert torch.autograd.gradcheck(lambda t: ag.irfft3(t, S), (G,), **kw)
    assert torch.autograd.gradcheck(lambda t: ag.grad(t, S), (x,), **kw)
    v = [torch.randn(n, generator=gen, device=dev, dtype=torch.float64, requires_grad=True)
         for _ in range(3)]
    assert torch.autograd.gradcheck(lambda a, b, c: ag.divergence(a, b, c, S), tuple(v), **kw)
    for f in (ag.laplace, ag.inv_laplace, ag.biharmonic, ag.inv_biharmonic):
        assert torch.autograd.gradcheck(lambda t: f(t, S), (x,), **kw), f.__name__


def _loss(u, g, S):
    from accfft_amd import autograd as ag
    r = ag.divergence(*ag.grad(ag.inv_laplace(u, S), S), S) - g
    return 0.5 * (r.double() ** 2).sum()


def test_loss_fd_and_closed_form_256(dev, orc):
    """L(u) = 1/2 |div(grad(invlap u)) - g|^2 at 256^3 FP32: autograd gradient
    vs central differences (eps = 1e-3, 64 random directions) and vs the
    closed form P(u - g) evaluated with the forward chain."""
    import torch
    from accfft_amd import autograd as ag
    n = (256, 256, 256)
    print(f"loss FD check {n}", flush=True)
    S = ag.Spectral(n, torch.float32, dev)
    gen = torch.Generator(device=dev).manual_seed(54)
    # the workload's inputs: u ~ N(0, 1), g band-limited with a |k|^-2 spectrum
    u = torch.tensor(fields.normal(n, fields.SEED0 + 3), dtype=torch.float32, device=dev,
                     requires_grad=True)
    g = torch.tensor(fields.band_limited(n, fields.SEED0 + 3, slope=-2.0), dtype=torch.float32,
                     device=dev)
    saved = []
    with torch.autograd.graph.saved_tensors_hooks(lambda t: saved.append(t) or t, lambda t: t):
        L = _loss(u, g, S)
    # the only saved tensor is the residual held by the squared-norm node
    assert len(saved) <= 2, len(saved)
    L.backward()
    grad = u.grad.detach()
    # closed form: P(u - g), P the projection (its FP64 value is pinned on a
    # 32^3 grid against the oracle below).  It is evaluated in FP64: applied
    # as div(grad(invlap(.))), the smooth target's large low modes leave
    # FP32 rounding noise in every mode of invlap(u - g) that the two
    # derivatives then amplify by |k|^2 (up to ~5e4), a 9e-4 error in FP32
    # (the backward applies the derivatives before invlap and has no such
    # amplification)
    S64 = ag.Spectral(n, torch.float64, dev)
    Pu = ag.divergence(*ag.grad(ag.inv_laplace((u.detach() - g).double(), S64), S64), S64)
    S64.release()
    assert rel_l2(grad.cpu().numpy(), Pu.cpu().numpy()) < 1e-5
    del Pu
    # central differences along 64 directions d ~ N(0, 1) per point (the
    # loss is quadratic, so the difference quotient is exact up to rounding);
    # the error is scaled by |dL/du| |d| (Cauchy-Schwarz bound of <dL/du, d>)
    eps = 1e-3
    worst = worst_rel = 0.0
    gnorm = grad.double().norm().item()
    with torch.no_grad():
        for i in range(64):
            d = torch.randn(n, generator=gen, device=dev)
            fd = (_loss(u + eps * d, g, S) - _loss(u - eps * d, g, S)).item() / (2 * eps)
            an = _dot(grad, d)
            worst = max(worst, abs(fd - an) / (gnorm * d.double().norm().item()))
            worst_rel = max(worst_rel, abs(fd - an) / abs(an))
    print(f"  FD over 64 directions: worst |fd-an|/(|g||d|) {worst:.2e}, "
          f"worst |fd-an|/|an| {worst_rel:.2e}")
    assert worst < 1e-5
    S.release()
    assert S.plan.workspace_bytes() == 0


def test_loss_closed_form_vs_oracle(dev, orc):
    """dL/du = P(u - g) against the FP64 oracle's projection on a 32^3 grid."""
    import torch
    from accfft_amd import autograd as ag
    n = (32, 32, 32)
    gen = np.random.default_rng(55)
    u0, g0 = gen.standard_normal(n), gen.standard_normal(n)
    u = torch.tensor(u0, device=dev, requires_grad=True)
    g = torch.tensor(g0, device=dev)
    _loss(u, g, None).backward()
    P = lambda a: orc.divergence(*orc.grad(orc.inv_laplace(a)))  # noqa: E731
    assert rel_l2(u.grad.cpu().numpy(), P(P(u0) - g0)) < 1e-12


def test_checkpoint_matches(dev):
    """torch.utils.checkpoint around the chain gives the identical gradient."""
    import torch
    import torch.utils.checkpoint as ckpt
    n = (64, 32, 48)
    gen = torch.Generator(device=dev).manual_seed(56)
    u = torch.randn(n, generator=gen, device=dev, requires_grad=True)
    g = torch.randn(n, generator=gen, device=dev)
    _loss(u, g, None).backward()
    ref = u.grad.clone()
    u.grad = None
    ckpt.checkpoint(_loss, u, g, None, use_reentrant=False).backward()
    assert torch.equal(u.grad, ref)
