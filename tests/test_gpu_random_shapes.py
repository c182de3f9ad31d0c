"""Seeded random shapes (every axis 1 ... 70, primes and odd lengths
included) through r2c, c2r and c2c in both precisions, and random pencil
grids over them, against the oracle: the edge cases a fixed list of shapes
can miss (axes of one point, Nyquist bins of odd lengths, ragged pencil
blocks)."""
import numpy as np
import pytest

from conftest import rel_l2, rng

pytestmark = pytest.mark.gpu

_R = np.random.default_rng(20261203)
SHAPES = [tuple(int(v) for v in _R.integers(1, 71, size=3)) for _ in range(24)]
SHAPES += [(1, 1, 1), (1, 5, 1), (7, 1, 2), (2, 3, 1), (1, 1, 9)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_random_shape_vs_oracle(dev, orc, shape, prec):
    import torch
    import accfft_amd as af
    print(f"random shape {shape} {prec}", flush=True)
    fp, rdt, cdt, tol = ((af.F64, torch.float64, torch.complex128, 1e-12) if prec == "f64"
                         else (af.F32, torch.float32, torch.complex64, 1e-5))
    n = list(shape)
    x = rng(sum(shape)).standard_normal(shape)
    xd = torch.tensor(x, dtype=rdt, device=dev)
    h = torch.empty((n[0], n[1], n[2] // 2 + 1), dtype=cdt, device=dev)
    p = af.Plan("r2c", n, xd, h, fp, af.ACCFFT_ESTIMATE)
    p.execute_r2c(xd, h)
    torch.cuda.synchronize()
    ref = orc.r2c(x)
    assert rel_l2(h.cpu().numpy(), ref) < tol
    back = torch.empty_like(xd)
    p.execute_c2r(torch.tensor(ref, dtype=cdt, device=dev), back)
    torch.cuda.synchronize()
    assert rel_l2(back.cpu().numpy(), orc.c2r(ref, n)) < tol
    p.destroy()
    c = x + 1j * rng(sum(shape) + 1).standard_normal(shape)
    cd = torch.tensor(c, dtype=cdt, device=dev)
    co = torch.empty_like(cd)
    q = af.Plan("c2c", n, cd, co, fp, af.ACCFFT_ESTIMATE)
    q.execute_c2c(af.ACCFFT_FORWARD, cd, co)
    torch.cuda.synchronize()
    assert rel_l2(co.cpu().numpy(), orc.c2c(c, -1)) < tol
    q.execute_c2c(af.ACCFFT_BACKWARD, co, cd)
    torch.cuda.synchronize()
    assert rel_l2(cd.cpu().numpy() / np.prod(n), c) < tol
    q.destroy()


@pytest.mark.parametrize("shape", SHAPES[:12] + SHAPES[-5:])
def test_random_shape_operators_fp64(dev, orc, shape):
    """laplace, inverse laplace, grad and divergence on the random shapes
    (odd lengths exercise the multiplier's Nyquist rule)."""
    import torch
    import accfft_amd as af
    print(f"random shape operators {shape}", flush=True)
    x = rng(sum(shape) + 2).standard_normal(shape)
    xd = torch.tensor(x, device=dev)
    p = af.Plan("r2c", list(shape), xd, None, af.F64, af.ACCFFT_ESTIMATE)
    out = torch.empty_like(xd)
    p.laplace(xd, out)
    torch.cuda.synchronize()
    assert rel_l2(out.cpu().numpy(), orc.laplace(x)) < 1e-12
    p.inv_laplace(xd, out)
    torch.cuda.synchronize()
    assert rel_l2(out.cpu().numpy(), orc.inv_laplace(x)) < 1e-12
    g = [torch.empty_like(xd) for _ in range(3)]
    p.grad(xd, *g)
    torch.cuda.synchronize()
    ref = orc.grad(x, 7)
    for d in range(3):
        assert rel_l2(g[d].cpu().numpy(), ref[d]) < 1e-12, d
    p.divergence(g[0], g[1], g[2], out)
    torch.cuda.synchronize()
    assert rel_l2(out.cpu().numpy(), orc.divergence(*[gi.cpu().numpy() for gi in g])) < 1e-12
    p.destroy()


_G = np.random.default_rng(20261204)
GRIDS = [(SHAPES[i], (int(_G.integers(1, 5)), int(_G.integers(1, 5)))) for i in range(10)]


@pytest.mark.parametrize("shape,grid", GRIDS)
def test_random_pencil_grid_vs_oracle(dev, orc, shape, grid):
    """Every rank of a random p0 x p1 grid over a random shape (one thread and
    stream per rank, the in-process all-to-all-v of test_gpu_pencil_share)."""
    import itertools

    import torch
    import accfft_amd as af
    from test_gpu_pencil_share import World, _hip, _run_ranks
    print(f"random pencil {shape} grid {grid}", flush=True)
    N0, N1, N2 = shape
    Nh = N2 // 2 + 1
    g = rng(sum(shape) + 3).standard_normal(shape)
    G = orc.r2c(g)
    totals, plans, ins, outs = {}, {}, {}, {}
    for c0, c1 in itertools.product(range(grid[0]), range(grid[1])):
        p = af.Plan("r2c_pencil", shape, None, None, af.F64, af.ACCFFT_ESTIMATE, grid=grid,
                    coords=(c0, c1))
        n0, n1, _ = p.isize
        ny, nz = p.osize[1], p.osize[2]
        totals[(c0, c1, 0, -1)] = (n0 * n1 * Nh * 16, n0 * N1 * nz * 16)
        totals[(c0, c1, 1, -1)] = (n0 * N1 * nz * 16, N0 * ny * nz * 16)
        totals[(c0, c1, 1, 1)] = totals[(c0, c1, 1, -1)][::-1]
        totals[(c0, c1, 0, 1)] = totals[(c0, c1, 0, -1)][::-1]
        xs, y1s = p.istart[0], p.istart[1]
        ins[(c0, c1)] = torch.tensor(np.ascontiguousarray(g[xs:xs + n0, y1s:y1s + n1]), device=dev)
        outs[(c0, c1)] = torch.empty(p.osize, dtype=torch.complex128, device=dev)
        plans[(c0, c1)] = p
    world = World(grid, _hip(), totals)
    for r, p in plans.items():
        p.set_exchange_v(world.hook(*r))
    _run_ranks({r: (lambda r=r: plans[r].execute_r2c(ins[r], outs[r])) for r in plans})
    for (c0, c1), p in plans.items():
        ys, zs = p.ostart[1], p.ostart[2]
        ref = G[:, ys:ys + p.osize[1], zs:zs + p.osize[2]]
        if ref.size:
            assert rel_l2(outs[(c0, c1)].cpu().numpy(), ref) < 1e-12, (c0, c1)
    backs = {r: torch.empty_like(ins[r]) for r in plans}
    _run_ranks({r: (lambda r=r: plans[r].execute_c2r(outs[r], backs[r])) for r in plans})
    for (c0, c1), p in plans.items():
        xs, y1s = p.istart[0], p.istart[1]
        want = g[xs:xs + p.isize[0], y1s:y1s + p.isize[1]] * np.prod(shape)
        if want.size:
            assert rel_l2(backs[(c0, c1)].cpu().numpy(), want) < 1e-12, (c0, c1)
    for p in plans.values():
        p.destroy()
