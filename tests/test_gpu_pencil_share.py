"""Pencil-share plans (accfft_amd_plan_dft_3d_r2c_pencil): every rank of a
p0 x p1 grid in one process, one thread per rank, each on its own stream;
the exchange_v hook copies the blocks between the ranks' device buffers
behind barriers (an all-to-all-v inside the process, what MPI_Alltoallv
does in the reference's pencil transposes, src/transpose_gpu.cpp:2058).

Each rank's forward output is checked against the oracle's global r2c
restricted to its block [N0][ys:ys+ny][zs:zs+nz], its inverse against the
oracle's global c2r restricted to [xs:xs+n0][y1s:y1s+n1]: ragged splits,
ranks with empty shares, communicators of one rank (no exchange issued),
and both forms of the copies (separate pack / unpack kernels, or carried
by the z and y passes when the y rows split evenly in power-of-two
blocks)."""
import ctypes
import itertools
import threading

import numpy as np
import pytest

from conftest import rel_l2, rng

pytestmark = pytest.mark.gpu


def _hip():
    """The HIP runtime this process already loaded (torch's)."""
    with open("/proc/self/maps") as f:
        for line in f:
            if "libamdhip64.so" in line:
                hip = ctypes.CDLL(line.split()[-1])
                hip.hipMemcpyAsync.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                               ctypes.c_int, ctypes.c_void_p]
                hip.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
                return hip
    raise RuntimeError("libamdhip64 not loaded")


def _prefix(v):
    return [int(x) for x in np.concatenate([[0], np.cumsum(v)[:-1]])] if len(v) else []


class World:
    """The ranks of one grid: barriers per communicator, the send buffers
    each rank posts, and the byte totals each exchange must move (computed
    here from the local sizes, independently of the plan)."""

    def __init__(self, grid, hip, totals):
        self.P0, self.P1 = grid
        self.hip = hip
        self.totals = totals  # (c0, c1, comm, direction) -> (send bytes, recv bytes)
        self.bars = {(0, c0): threading.Barrier(self.P1, timeout=60) for c0 in range(self.P0)}
        self.bars.update({(1, c1): threading.Barrier(self.P0, timeout=60) for c1 in range(self.P1)})
        self.posted = {}
        self.calls = []

    def hook(self, c0, c1):
        def fn(comm, send, sb, so, recv, rb, ro, direction, stream):
            key = (comm, c0 if comm == 0 else c1)
            me = c1 if comm == 0 else c0
            P = self.P1 if comm == 0 else self.P0
            bar = self.bars[key]
            try:
                self.calls.append((c0, c1, comm, direction))
                assert len(sb) == len(rb) == P
                assert so == _prefix(sb) and ro == _prefix(rb), (so, sb, ro, rb)
                assert (sum(sb), sum(rb)) == self.totals[(c0, c1, comm, direction)]
                assert self.hip.hipStreamSynchronize(stream) == 0
                self.posted[(key, me)] = (send, sb, so)
                bar.wait()
                for q in range(P):
                    qsend, qsb, qso = self.posted[(key, q)]
                    nb = qsb[me]
                    assert nb == rb[q], (comm, q, me, nb, rb[q])
                    if nb:
                        assert self.hip.hipMemcpyAsync(recv + ro[q], qsend + qso[me], nb, 3,
                                                       stream) == 0
                assert self.hip.hipStreamSynchronize(stream) == 0
                bar.wait()  # every peer has finished reading this rank's send buffer
                return 0
            except Exception:
                bar.abort()
                raise
        return fn


def _run_ranks(work):
    """Runs work(rank) for every rank in its own thread on its own stream,
    after everything queued so far (inputs written on the default stream)."""
    import torch
    torch.cuda.synchronize()
    errs = []

    def body(rank, fn):
        try:
            s = torch.cuda.Stream()
            with torch.cuda.stream(s):
                fn()
                s.synchronize()
        except Exception as e:  # noqa: BLE001 - reported below
            errs.append((rank, repr(e)))

    th = [threading.Thread(target=body, args=(r, f)) for r, f in work.items()]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=180)
    assert not any(t.is_alive() for t in th), "a rank did not finish"
    assert not errs, errs


CASES = [((8, 12, 10), (2, 3), "f64"), ((9, 10, 7), (4, 2), "f64"), ((6, 8, 16), (1, 4), "f64"),
         ((10, 6, 8), (3, 1), "f64"), ((8, 12, 6), (2, 3), "f64"), ((7, 5, 12), (3, 5), "f64"),
         ((16, 16, 18), (4, 4), "f64"), ((32, 24, 40), (2, 2), "f64"), ((5, 4, 3), (1, 1), "f64"),
         ((64, 48, 80), (2, 3), "f32"), ((128, 96, 160), (4, 2), "f32"),
         # y rows split in equal power-of-two blocks: the copies ride on the
         # z and y passes (ragged z column blocks, an empty z block, one-row
         # y blocks with empty x ranks)
         ((8, 16, 20), (2, 4), "f64"), ((16, 64, 36), (4, 4), "f64"), ((9, 12, 6), (3, 3), "f64"),
         ((7, 8, 10), (8, 8), "f64"), ((64, 64, 62), (2, 4), "f32")]


@pytest.mark.parametrize("N,grid,prec", CASES)
def test_pencil_share_vs_oracle(dev, orc, N, grid, prec):
    import torch
    import accfft_amd as af
    print(f"pencil share {N} grid {grid} {prec}", flush=True)
    fp = af.F64 if prec == "f64" else af.F32
    rdt = torch.float64 if prec == "f64" else torch.float32
    cdt = torch.complex128 if prec == "f64" else torch.complex64
    cs = 16 if prec == "f64" else 8
    tol = 1e-12 if prec == "f64" else 1e-5
    P0, P1 = grid
    N0, N1, N2 = N
    Nh = N2 // 2 + 1
    g = rng(81).standard_normal(N)
    G = orc.r2c(g)
    ginv = orc.c2r(G, list(N))
    hip = _hip()
    totals, plans, ins, outs, backs = {}, {}, {}, {}, {}
    for c0, c1 in itertools.product(range(P0), range(P1)):
        _, isz, ist, osz, ost = af.local_size_dft_r2c(N, fp, grid, (c0, c1))
        n0, n1, ny, nz = isz[0], isz[1], osz[1], osz[2]
        totals[(c0, c1, 0, -1)] = (n0 * n1 * Nh * cs, n0 * N1 * nz * cs)
        totals[(c0, c1, 1, -1)] = (n0 * N1 * nz * cs, N0 * ny * nz * cs)
        totals[(c0, c1, 1, 1)] = totals[(c0, c1, 1, -1)][::-1]
        totals[(c0, c1, 0, 1)] = totals[(c0, c1, 0, -1)][::-1]
        x = np.ascontiguousarray(g[ist[0]:ist[0] + n0, ist[1]:ist[1] + n1, :])
        ins[(c0, c1)] = torch.tensor(x, dtype=rdt, device=dev)
        outs[(c0, c1)] = torch.empty((N0, ny, nz), dtype=cdt, device=dev)
        backs[(c0, c1)] = torch.empty((n0, n1, N2), dtype=rdt, device=dev)
        p = af.Plan("r2c_pencil", N, None, None, fp, af.ACCFFT_ESTIMATE, grid=grid,
                    coords=(c0, c1))
        assert p.isize == tuple(isz) and p.osize == tuple(osz) and p.ostart == tuple(ost)
        plans[(c0, c1)] = p
    world = World(grid, hip, totals)
    for r, p in plans.items():
        p.set_exchange_v(world.hook(*r))
    timers = {r: [0.0] * 5 for r in plans}
    _run_ranks({r: (lambda r=r: plans[r].execute_r2c(ins[r], outs[r], timer=timers[r]))
                for r in plans})
    for (c0, c1), p in plans.items():
        ys, zs = p.ostart[1], p.ostart[2]
        ny, nz = p.osize[1], p.osize[2]
        ref = G[:, ys:ys + ny, zs:zs + nz]
        if ref.size:
            assert rel_l2(outs[(c0, c1)].cpu().numpy(), ref) < tol, (c0, c1)
        if P0 > 1 and P1 > 1:
            assert timers[(c0, c1)][2] > 0 and timers[(c0, c1)][4] > 0, timers[(c0, c1)]
    # one-rank communicators issue no exchange (transpose_gpu.cpp:1814-1816)
    comms = {c for (_, _, c, _) in world.calls}
    assert comms == {c for c, P in ((0, P1), (1, P0)) if P > 1}
    _run_ranks({r: (lambda r=r: plans[r].execute_c2r(outs[r], backs[r])) for r in plans})
    for (c0, c1), p in plans.items():
        xs, y1s = p.istart[0], p.istart[1]
        n0, n1 = p.isize[0], p.isize[1]
        ref = ginv[xs:xs + n0, y1s:y1s + n1, :]
        if ref.size:
            got = backs[(c0, c1)].cpu().numpy()
            assert rel_l2(got, ref) < tol, (c0, c1)
            assert rel_l2(got / np.prod(N), g[xs:xs + n0, y1s:y1s + n1, :]) < tol
    for p in plans.values():
        p.destroy()


def test_pencil_identity_exchange(dev):
    """Without a hook each exchange is the identity (packed blocks taken as
    received), allowed where every block sent has the size of the block
    received: same bits as a hook that copies this rank's own blocks.  A
    ragged split without a hook is refused."""
    import torch
    import accfft_amd as af
    N, grid, c = (8, 8, 14), (2, 2), (1, 0)
    hip = _hip()
    p = af.Plan("r2c_pencil", N, None, None, af.F64, af.ACCFFT_ESTIMATE, grid=grid, coords=c)
    n0, n1, _ = p.isize
    x = torch.randn((n0, n1, N[2]), dtype=torch.float64, device=dev)
    o0 = torch.empty(p.osize, dtype=torch.complex128, device=dev)
    o1 = torch.empty_like(o0)
    p.execute_r2c(x, o0)
    b0 = torch.empty_like(x)
    p.execute_c2r(o0.clone(), b0)
    calls = []

    def self_copy(comm, send, sb, so, recv, rb, ro, direction, stream):
        calls.append((comm, direction))
        assert sb == rb and so == ro
        return 0 if hip.hipMemcpyAsync(recv, send, sum(sb), 3, stream) == 0 else 5

    p.set_exchange_v(self_copy)
    p.execute_r2c(x, o1)
    b1 = torch.empty_like(x)
    p.execute_c2r(o1.clone(), b1)
    torch.cuda.synchronize()
    assert torch.equal(o0, o1) and torch.equal(b0, b1)
    assert calls == [(0, -1), (1, -1), (1, 1), (0, 1)]
    p.destroy()
    q = af.Plan("r2c_pencil", (9, 10, 7), None, None, af.F64, af.ACCFFT_ESTIMATE, grid=(4, 2),
                coords=(0, 0))
    xi = torch.zeros(q.isize, dtype=torch.float64, device=dev)
    oi = torch.empty(q.osize, dtype=torch.complex128, device=dev)
    with pytest.raises(af.AccfftError, match="unsupported"):
        q.execute_r2c(xi, oi)
    q.destroy()


def test_pencil_plan_arguments(dev):
    import accfft_amd as af
    with pytest.raises(af.AccfftError, match="invalid"):
        af.Plan("r2c_pencil", (8, 8, 8), None, None, af.F64, grid=(2, 2), coords=(2, 0))
    with pytest.raises(af.AccfftError, match="unsupported"):
        af.Plan("r2c_pencil", (8, 8, 8), None, None, af.F64, grid=(65, 1), coords=(0, 0))
    with pytest.raises(af.AccfftError, match="unsupported"):  # lines longer than one LDS tile
        af.Plan("r2c_pencil", (8, 8, 8192), None, None, af.F64, grid=(2, 2), coords=(0, 0))
    p = af.Plan("r2c_pencil", (8, 8, 8), None, None, af.F64, grid=(2, 2), coords=(1, 1))
    with pytest.raises(af.AccfftError, match="wrong kind"):
        p.set_exchange(None)
    p.destroy()


def test_pencil_share_2048_fp32_full_grid(dev):
    """2048^3 FP32 on a 4 x 2 grid, all eight ranks at their real local
    shapes ([512][1024][2048] real -> [2048][512][513 | 512] complex: the
    row transposes are ragged) on the one GPU (about 180 GB).  Forward of
    cos(2 pi (kx x + ky y + kz z) / 2048): the one non-zero bin of the half
    spectrum, N^3 / 2 at (kx, ky, kz), lies in the block of rank (2, 1) and
    every other bin of every rank is zero; round trip of N(0,1) blocks."""
    import torch
    import accfft_amd as af
    N, grid = (2048, 2048, 2048), (4, 2)
    kx, ky, kz = 3, 1500, 700
    print(f"pencil share {N} grid {grid} FP32, all ranks", flush=True)
    torch.cuda.empty_cache()  # the plans' workspaces come from hipMalloc, beside torch's cache
    hip = _hip()
    cs = 8
    Nh = N[2] // 2 + 1
    totals, plans, ins, outs = {}, {}, {}, {}
    for c0, c1 in itertools.product(range(grid[0]), range(grid[1])):
        p = af.Plan("r2c_pencil", N, None, None, af.F32, af.ACCFFT_ESTIMATE, grid=grid,
                    coords=(c0, c1))
        n0, n1, _ = p.isize
        ny, nz = p.osize[1], p.osize[2]
        totals[(c0, c1, 0, -1)] = (n0 * n1 * Nh * cs, n0 * N[1] * nz * cs)
        totals[(c0, c1, 1, -1)] = (n0 * N[1] * nz * cs, N[0] * ny * nz * cs)
        totals[(c0, c1, 1, 1)] = totals[(c0, c1, 1, -1)][::-1]
        totals[(c0, c1, 0, 1)] = totals[(c0, c1, 0, -1)][::-1]
        xs, ys1 = p.istart[0], p.istart[1]
        t = 2 * np.pi / N[0]
        x = torch.arange(xs, xs + n0, device=dev, dtype=torch.float64)
        y = torch.arange(ys1, ys1 + n1, device=dev, dtype=torch.float64)
        z = torch.arange(N[2], device=dev, dtype=torch.float64)
        a = ((kx * x).view(-1, 1) + (ky * y).view(1, -1)) % N[0] * t
        b = (kz * z) % N[2] * t
        u = torch.cos(a).float().unsqueeze(-1) * torch.cos(b).float()
        u.sub_(torch.sin(a).float().unsqueeze(-1) * torch.sin(b).float())
        ins[(c0, c1)] = u
        outs[(c0, c1)] = torch.empty(p.osize, dtype=torch.complex64, device=dev)
        plans[(c0, c1)] = p
    try:
        _full_grid_checks(dev, N, grid, (kx, ky, kz), hip, totals, plans, ins, outs)
    finally:  # release the ~180 GB also when a check fails
        for p in plans.values():
            p.destroy()
        ins.clear()
        outs.clear()
        torch.cuda.empty_cache()


def _full_grid_checks(dev, N, grid, k, hip, totals, plans, ins, outs):
    import torch
    kx, ky, kz = k
    world = World(grid, hip, totals)
    for r, p in plans.items():
        p.set_exchange_v(world.hook(*r))
    _run_ranks({r: (lambda r=r: plans[r].execute_r2c(ins[r], outs[r])) for r in plans})
    peak = float(N[0]) * N[1] * N[2] / 2
    hit = []
    rest2 = 0.0
    for (c0, c1), p in plans.items():
        o = outs[(c0, c1)]
        ys, zs = p.ostart[1], p.ostart[2]
        ny, nz = p.osize[1], p.osize[2]
        if ys <= ky < ys + ny and zs <= kz < zs + nz:
            v = o[kx, ky - ys, kz - zs].item()
            assert abs(v - peak) / peak < 1e-5, (c0, c1, v)
            o[kx, ky - ys, kz - zs] = 0
            hit.append((c0, c1))
        rest2 += (o.abs().double().norm() ** 2).item()
    assert hit == [(2, 1)], hit
    rest = rest2 ** 0.5 / peak
    print(f"  closed form: rest of the spectrum {rest:.2e} of the peak", flush=True)
    assert rest < 1e-5
    # round trip of N(0,1) blocks, each rank's result written over its input
    for i, r in enumerate(plans):
        g = torch.Generator(device=dev).manual_seed(90 + i)
        ins[r].copy_(torch.randn(ins[r].shape, generator=g, device=dev))
    _run_ranks({r: (lambda r=r: plans[r].execute_r2c(ins[r], outs[r])) for r in plans})
    _run_ranks({r: (lambda r=r: plans[r].execute_c2r(outs[r], ins[r])) for r in plans})
    scale = float(np.prod(N))
    for i, r in enumerate(plans):
        g = torch.Generator(device=dev).manual_seed(90 + i)
        want = torch.randn(ins[r].shape, generator=g, device=dev)
        err = ((ins[r] / scale - want).double().norm() / want.double().norm()).item()
        assert err < 1e-5, (r, err)
        del want
