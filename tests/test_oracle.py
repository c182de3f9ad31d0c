"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

Each check is chosen so that a plausible slip in the oracle (wrong sign,
dropped twiddle, transposed axes, wrong Nyquist handling, wrong stage order)
fails at least one of them:
  * golden vectors from numpy.fft / pocketfft (tests/golden/fixtures.npz)
  * brute-force O(n^2) DFT in long double for every length 1..64 and primes
  * closed forms: sin x + sin y + sin z, delta, constant
  * invariants: linearity, Parseval with Hermitian weights, round trip
  * per-axis transforms against numpy.fft.rfft along that axis
  * the multiplier tables against hand-computed wavenumbers, including the
    odd-N Nyquist quirk of src/operators_cuda.cu:34-37
  * operators on analytic fields (steps/step5/step5_gpu.cpp:20-133)
"""
import os

import numpy as np
import pytest

from conftest import rel_l2, rng

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fixtures.npz")


@pytest.fixture(scope="module")
def fx():
    with np.load(FIX, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


# ------------------------------------------------------------------ golden vectors
@pytest.mark.parametrize("key", ["8x6x5", "12x10x9", "16x16x16", "30x20x17"])
def test_r2c_golden(orc, fx, key):
    a = fx[f"r2c_{key}_in"]
    assert rel_l2(orc.r2c(a), fx[f"r2c_{key}_out"]) < 1e-14
    back = orc.c2r(fx[f"r2c_{key}_out"], a.shape)
    assert rel_l2(back / a.size, a) < 1e-14


def test_c2c_golden(orc, fx):
    c = fx["c2c_6x10x7_in"]
    assert rel_l2(orc.c2c(c, -1), fx["c2c_6x10x7_out"]) < 1e-14
    assert rel_l2(orc.c2c(fx["c2c_6x10x7_out"], +1) / c.size, c) < 1e-14


@pytest.mark.parametrize("key", ["8x6x10", "5x7x9"])
def test_c2r_nonhermitian_golden(orc, fx, key):
    h = fx[f"irfft_{key}_in"]
    shp = tuple(int(v) for v in key.split("x"))
    assert rel_l2(orc.c2r(h, shp), fx[f"irfft_{key}_out"]) < 1e-14


def test_64_sampled(orc, fx):
    a = np.random.default_rng(int(fx["s64_seed"])).standard_normal((64, 64, 64))
    s = orc.r2c(a)
    np.testing.assert_allclose(s.ravel()[fx["s64_idx"]], fx["s64_vals"], rtol=0, atol=1e-10)
    assert abs(np.linalg.norm(s) - float(fx["s64_norm"])) / float(fx["s64_norm"]) < 1e-14
    assert abs(s.sum() - complex(fx["s64_sum"])) < 1e-8


# ------------------------------------------------------------------ 1-D against brute force
@pytest.mark.parametrize("n", list(range(1, 65)) + [97, 101, 128, 243, 256, 360, 512, 625])
def test_dft1d_vs_naive(orc, n):
    x = rng(n).standard_normal(n) + 1j * rng(n + 1).standard_normal(n)
    for sign in (-1, 1):
        a = orc.dft1d(x, sign)
        b = orc.dft1d_naive(x, sign)
        assert rel_l2(a, b) < 2e-15 * max(1.0, np.log2(n))


# ------------------------------------------------------------------ closed forms
@pytest.mark.parametrize("N", [8, 9, 16, 30])
def test_sin_closed_form(orc, N):
    g = 2 * np.pi * np.arange(N) / N
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    s = orc.r2c(np.sin(X) + np.sin(Y) + np.sin(Z))
    ref = np.zeros_like(s)
    c = -0.5j * N ** 3
    ref[1, 0, 0] = c
    ref[0, 1, 0] = c
    ref[0, 0, 1] = c
    ref[N - 1, 0, 0] = -c
    ref[0, N - 1, 0] = -c
    np.testing.assert_allclose(s, ref, atol=1e-10 * N ** 3)


def test_delta_and_constant(orc):
    n = (6, 8, 10)
    d = np.zeros(n)
    d[0, 0, 0] = 1
    np.testing.assert_allclose(orc.r2c(d), np.ones((6, 8, 6)), atol=1e-15)
    c = np.full(n, 2.5)
    s = orc.r2c(c)
    ref = np.zeros_like(s)
    ref[0, 0, 0] = 2.5 * np.prod(n)
    np.testing.assert_allclose(s, ref, atol=1e-12)
    # shifted delta -> linear phase
    d2 = np.zeros(n)
    d2[1, 2, 3] = 1
    k0, k1, k2 = np.meshgrid(np.arange(6), np.arange(8), np.arange(6), indexing="ij")
    ph = np.exp(-2j * np.pi * (k0 * 1 / 6 + k1 * 2 / 8 + k2 * 3 / 10))
    np.testing.assert_allclose(orc.r2c(d2), ph, atol=1e-14)


def test_linearity_parseval_roundtrip(orc):
    n = (12, 9, 14)
    a, b = rng(1).standard_normal(n), rng(2).standard_normal(n)
    sa, sb = orc.r2c(a), orc.r2c(b)
    assert rel_l2(orc.r2c(2 * a - 3 * b), 2 * sa - 3 * sb) < 1e-14
    # Parseval with Hermitian weights: bins 0 and N/2 along z count once
    w = np.full(n[2] // 2 + 1, 2.0)
    w[0] = 1
    if n[2] % 2 == 0:
        w[-1] = 1
    lhs = np.sum(w * np.abs(sa) ** 2) / np.prod(n)
    assert abs(lhs - np.sum(a * a)) / np.sum(a * a) < 1e-13
    assert rel_l2(orc.c2r(sa, n) / np.prod(n), a) < 1e-14


@pytest.mark.parametrize("shape", [(8, 6, 5), (7, 9, 11), (16, 12, 10)])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_axis_transforms(orc, shape, axis):
    a = rng(3).standard_normal(shape)
    h = orc.r2c_axis(a, axis)
    ref = np.moveaxis(np.fft.rfft(np.moveaxis(a, axis, -1)), -1, axis)
    assert rel_l2(h, ref) < 1e-14
    assert rel_l2(orc.c2r_axis(h, shape, axis) / shape[axis], a) < 1e-14


def test_xyz_mask(orc):
    n = (8, 6, 10)
    a = rng(4).standard_normal(n)
    # z and x only (used by divergence_slow): rfft along z then fft along x
    ref = np.fft.fft(np.fft.rfft(a, axis=2), axis=0)
    assert rel_l2(orc.r2c(a, xyz=0b101), ref) < 1e-14
    assert rel_l2(orc.c2r(ref, n, xyz=0b101) / (8 * 10), a) < 1e-14


# ------------------------------------------------------------------ multipliers
def _w(X, N):
    """Hand-written wavenumber table with the Nyquist filter."""
    w = X if X <= N // 2 else X - N
    return 0 if X == N // 2 else w


def test_grad_symbol_even_and_odd(orc):
    for N in (8, 7):
        h = np.ones((N, 1, 1), complex)
        out = orc.mult(orc.GRAD_X, (N, 1, 1), h, scale=1.0)
        expect = np.array([1j * _w(X, N) for X in range(N)])
        np.testing.assert_allclose(out.ravel(), expect, atol=0)
    # explicit odd quirk: N = 7 zeroes +3 but keeps -3
    out = orc.mult(orc.GRAD_X, (7, 1, 1), np.ones((7, 1, 1), complex)).ravel()
    assert out[3] == 0 and out[4] == -3j


def test_laplace_and_inverse_symbols(orc):
    N = (8, 6, 10)
    h = np.ones((8, 6, 6), complex)
    lap = orc.mult(orc.LAPLACE, N, h).real
    inv = orc.mult(orc.INV_LAPLACE, N, h).real
    bih = orc.mult(orc.BIHARMONIC, N, h).real
    for (i, j, k) in [(0, 0, 0), (1, 2, 3), (4, 3, 5), (7, 5, 2), (3, 0, 0)]:
        wave = -(_w(i, 8) ** 2 + _w(j, 6) ** 2 + _w(k, 10) ** 2)
        assert lap[i, j, k] == wave
        assert bih[i, j, k] == wave * wave
        assert inv[i, j, k] == (0 if wave == 0 else 1.0 / wave)
    # zero set of the inverse: every w_d == 0 -> k_d in {0, N_d/2}: 8 modes
    assert int(np.sum(inv == 0)) == 8


# ------------------------------------------------------------------ operators on analytic fields
def _grid(n):
    g = [2 * np.pi * np.arange(m) / m for m in n]
    return np.meshgrid(*g, indexing="ij")


def test_operators_single_mode(orc):
    n = (16, 12, 20)
    X, Y, Z = _grid(n)
    f = np.sin(X) * np.sin(Y) * np.sin(Z)
    gx, gy, gz = orc.grad(f)
    np.testing.assert_allclose(gx, np.cos(X) * np.sin(Y) * np.sin(Z), atol=1e-13)
    np.testing.assert_allclose(gy, np.sin(X) * np.cos(Y) * np.sin(Z), atol=1e-13)
    np.testing.assert_allclose(gz, np.sin(X) * np.sin(Y) * np.cos(Z), atol=1e-13)
    gs = orc.grad(f, slow=True)
    for a, b in zip(gs, (gx, gy, gz)):
        np.testing.assert_allclose(a, b, atol=1e-13)
    np.testing.assert_allclose(orc.laplace(f), -3 * f, atol=1e-13)
    np.testing.assert_allclose(orc.biharmonic(f), 9 * f, atol=1e-12)
    np.testing.assert_allclose(orc.inv_laplace(f), -f / 3, atol=1e-14)
    np.testing.assert_allclose(orc.divergence(gx, gy, gz), -3 * f, atol=1e-13)
    np.testing.assert_allclose(orc.divergence(gx, gy, gz, slow=True), -3 * f, atol=1e-13)


def test_gaussian_step1_field(orc):
    """exp(-4|x-pi|^2), the field of steps/step1/step1_gpu.cpp:135-145; grad
    and laplace against closed forms with the reference's rel-L1 < 1e-10."""
    n = (64, 64, 64)
    X, Y, Z = _grid(n)
    r2 = (X - np.pi) ** 2 + (Y - np.pi) ** 2 + (Z - np.pi) ** 2
    f = np.exp(-4 * r2)
    gx, gy, gz = orc.grad(f)
    ex = -8 * (X - np.pi) * f
    assert np.abs(gx - ex).sum() / np.abs(ex).sum() < 1e-10
    lap = (64 * r2 - 24) * f
    assert np.abs(orc.laplace(f) - lap).sum() / np.abs(lap).sum() < 1e-10


def test_div_grad_equals_laplace_even(orc):
    n = (16, 8, 12)
    a = rng(5).standard_normal(n)
    g = orc.grad(a)
    assert rel_l2(orc.divergence(*g), orc.laplace(a)) < 1e-13


def test_div_grad_invlap_projection(orc):
    """div(grad(invlap(u))) = P u, P zeroing the modes with every w_d = 0."""
    n = (8, 6, 10)
    u = rng(6).standard_normal(n)
    lhs = orc.divergence(*orc.grad(orc.inv_laplace(u)))
    h = orc.r2c(u)
    mask = np.ones(h.shape)
    for i in (0, 4):
        for j in (0, 3):
            for k in (0, 5):
                mask[i, j, k] = 0
    assert rel_l2(lhs, orc.c2r(h * mask, n) / np.prod(n)) < 1e-13


# ---------------------------------------------------------------- adjoints
# Each adjoint formula is checked against the explicit transpose of the map's
# matrix, assembled column by column from the forward oracle on tiny grids
# (odd and even lengths on every axis, so DC-only and DC+Nyquist z planes).
ADJ_SHAPES = [(4, 3, 6), (3, 4, 5), (2, 5, 4)]


def _basis(n):
    N = int(np.prod(n))
    for j in range(N):
        e = np.zeros(N)
        e[j] = 1.0
        yield j, e.reshape(n)


@pytest.mark.parametrize("n", ADJ_SHAPES)
def test_adjoint_r2c_is_matrix_transpose(orc, n):
    N = int(np.prod(n))
    M = np.stack([orc.r2c(e).ravel() for _, e in _basis(n)], axis=1)  # complex [K][N]
    g = rng(40)
    G = (g.standard_normal(M.shape[0]) + 1j * g.standard_normal(M.shape[0]))
    expect = (M.real.T @ G.real + M.imag.T @ G.imag).reshape(n)
    assert rel_l2(orc.adjoint_r2c(G.reshape(n[0], n[1], n[2] // 2 + 1), n), expect) < 1e-13
    assert M.shape == (n[0] * n[1] * (n[2] // 2 + 1), N)


@pytest.mark.parametrize("n", ADJ_SHAPES)
def test_adjoint_c2r_is_matrix_transpose(orc, n):
    hs = (n[0], n[1], n[2] // 2 + 1)
    K = int(np.prod(hs))
    cols_re, cols_im = [], []
    for k in range(K):
        e = np.zeros(K, np.complex128)
        e[k] = 1.0
        cols_re.append(orc.c2r(e.reshape(hs), n).ravel())
        e[k] = 1j
        cols_im.append(orc.c2r(e.reshape(hs), n).ravel())
    y = rng(41).standard_normal(int(np.prod(n)))
    expect = (np.array(cols_re) @ y + 1j * (np.array(cols_im) @ y)).reshape(hs)
    assert rel_l2(orc.adjoint_c2r(y.reshape(n)), expect) < 1e-13


@pytest.mark.parametrize("n", ADJ_SHAPES)
def test_adjoint_grad_div_are_matrix_transposes(orc, n):
    cols = [[], [], []]
    for _, e in _basis(n):
        for d, gd in enumerate(orc.grad(e)):
            cols[d].append(gd.ravel())
    Md = [np.array(c).T for c in cols]  # [N][N] per component
    g = rng(42)
    gs = [g.standard_normal(n) for _ in range(3)]
    expect = sum(Md[d].T @ gs[d].ravel() for d in range(3)).reshape(n)
    assert rel_l2(orc.adjoint_grad(*gs), expect) < 1e-12
    # divergence is the row block [Dx Dy Dz]; its transpose is (Dx^T v, Dy^T v, Dz^T v)
    v = g.standard_normal(n)
    for d, o in enumerate(orc.adjoint_divergence(v)):
        assert rel_l2(o, (Md[d].T @ v.ravel()).reshape(n)) < 1e-12
    # and the divergence oracle is that row block
    assert rel_l2(orc.divergence(*gs), sum(Md[d] @ gs[d].ravel() for d in range(3)).reshape(n)) < 1e-12


@pytest.mark.parametrize("op", ["laplace", "inv_laplace", "biharmonic", "inv_biharmonic"])
@pytest.mark.parametrize("n", ADJ_SHAPES)
def test_spectral_ops_are_symmetric(orc, op, n):
    M = np.array([getattr(orc, op)(e).ravel() for _, e in _basis(n)]).T
    assert np.abs(M - M.T).max() <= 1e-12 * max(1.0, np.abs(M).max())


# ---------------------------------------------------------------- slab share
def test_slab_one_rank_is_plain_r2c(orc):
    a = rng(60).standard_normal((6, 8, 10))
    assert rel_l2(orc.slab_r2c(a, 1), orc.r2c(a)) < 1e-14
    h = orc.r2c(a)
    assert rel_l2(orc.slab_c2r(h, 10, 1), orc.c2r(h, (6, 8, 10))) < 1e-14


@pytest.mark.parametrize("shape,P", [((8, 16, 10), 4), ((12, 6, 9), 3), ((16, 8, 8), 8)])
def test_slab_round_trip_and_blocks(orc, shape, P):
    N0, N1, N2 = shape
    n0 = N0 // P
    a = rng(61).standard_normal((n0, N1, N2))
    h = orc.slab_r2c(a, P)
    assert h.shape == (N0, N1 // P, N2 // 2 + 1)
    assert rel_l2(orc.slab_c2r(h, N2, P), N0 * N1 * N2 * a) < 1e-13
    # block q of the received array (before the x transform) is the y-range
    # q of the z/y-transformed slab: check through the x-inverse
    zy = np.fft.fft(np.fft.rfft(a, axis=2), axis=1)
    recv = np.fft.ifft(h, axis=0)
    n1 = N1 // P
    for q in range(P):
        assert rel_l2(recv[q * n0:(q + 1) * n0], zy[:, q * n1:(q + 1) * n1]) < 1e-13
