"""CPU tests of the C ABI boundary (no GPU compute): exported symbols, the
local-size rules of src/accfft_common.cpp:49-256, and the error paths that
the host can force without a GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "accfft_amd.h")


@pytest.fixture(scope="module")
def af():
    import accfft_amd
    if not os.path.exists(accfft_amd.LIB_PATH):
        from accfft_amd.build import build
        build(verbose=False)
    accfft_amd.lib()
    return accfft_amd


def declared_symbols():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(accfft_amd_\w+)\s*\(", txt)))


def test_every_declared_symbol_is_exported(af):
    L = af.lib()
    syms = declared_symbols()
    assert len(syms) >= 30
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing


# Values of alloc_max for a 1x1 grid, bytes (from the reference's rules,
# worked out by hand: N0*N1*(N2/2+1)*2*sizeof(T) dominates for cubes).
ALLOC = [
    ((64, 64, 64), 8, 2_162_688),
    ((256, 256, 256), 8, 135_266_304),
    ((256, 256, 256), 4, 67_633_152),
    ((512, 512, 512), 8, 1_077_936_128),
    ((512, 512, 512), 4, 538_968_064),
    ((1024, 1024, 1024), 4, 4_303_355_904),  # > 2^31: the reference's int overflows
]


@pytest.mark.parametrize("n,elem,expect", ALLOC)
def test_alloc_max_r2c(af, orc, n, elem, expect):
    fp = af.F64 if elem == 8 else af.F32
    amax, isize, istart, osize, ostart = af.local_size_dft_r2c(n, fp)
    assert amax == expect
    assert isize == tuple(n) and istart == (0, 0, 0)
    assert osize == (n[0], n[1], n[2] // 2 + 1) and ostart == (0, 0, 0)
    assert orc.local_size_r2c(n, elem)[0] == expect


def test_alloc_max_c2c(af):
    amax, isize, _, osize, _ = af.local_size_dft_c2c((512, 512, 512), af.F32)
    assert amax == 1_073_741_824
    assert osize == (512, 512, 512)


SHAPES = [(8, 6, 5), (17, 9, 4), (64, 64, 64), (30, 20, 17), (2048, 2048, 2048), (5, 300, 12)]
GRIDS = [(1, 1), (2, 1), (8, 1), (2, 3), (4, 4), (3, 5)]


@pytest.mark.parametrize("n", SHAPES)
@pytest.mark.parametrize("grid", GRIDS)
def test_local_size_decomposition(af, orc, n, grid):
    """Over all ranks of a p0 x p1 grid the input blocks tile [N0][N1][N2]
    and the output blocks tile the transposed spectrum [N1][Nh][N0]-ordered
    osize exactly once (the decomposition semantics of dfft_get_local_size_t),
    and the C ABI agrees with the oracle's restatement."""
    p0, p1 = grid
    nh = n[2] // 2 + 1
    cover_in = np.zeros((n[0], n[1]), int)
    cover_out = np.zeros((n[1], nh), int)
    for c0 in range(p0):
        for c1 in range(p1):
            a = af.local_size_dft_r2c(n, af.F64, grid, (c0, c1))
            b = orc.local_size_r2c(n, 8, grid, (c0, c1))
            assert a == b
            _, isz, ist, osz, ost = a
            assert isz[2] == n[2] and ist[2] == 0
            cover_in[ist[0]:ist[0] + isz[0], ist[1]:ist[1] + isz[1]] += 1
            # output: x is whole (osize[0] = N0), y split over p0, z split over p1
            assert osz[0] == n[0] and ost[0] == 0
            cover_out[ost[1]:ost[1] + osz[1], ost[2]:ost[2] + osz[2]] += 1
    assert (cover_in == 1).all()
    assert (cover_out == 1).all()


def test_local_size_c2c_matches_oracle(af, orc):
    for n in SHAPES:
        for grid in GRIDS:
            for c0 in range(grid[0]):
                for c1 in range(grid[1]):
                    assert af.local_size_dft_c2c(n, af.F32, grid, (c0, c1)) == \
                        orc.local_size_c2c(n, 4, grid, (c0, c1))


# ------------------------------------------------------------------ host-forced errors
def test_errors_without_gpu(af):
    L = af.lib()
    i3 = ctypes.c_int * 3
    h = ctypes.c_void_p()
    # invalid dims are rejected before any HIP call
    st = L.accfft_amd_plan_dft_3d_r2c(i3(0, 4, 4), None, None, af.F64, 0, ctypes.byref(h))
    assert st == 1 and not h.value
    st = L.accfft_amd_plan_dft_3d_r2c(i3(4, -1, 4), None, None, af.F64, 0, ctypes.byref(h))
    assert st == 1
    # bad precision
    st = L.accfft_amd_plan_dft_3d_c2c(i3(4, 4, 4), None, None, 7, 0, ctypes.byref(h))
    assert st == 1
    # a slab share whose lines exceed one LDS tile is unsupported (FP64: > 5120;
    # 1 x 1 plans run such lines as global radix passes)
    L.accfft_amd_plan_dft_3d_r2c_slab.argtypes = [i3, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                                  ctypes.c_int, ctypes.c_uint, ctypes.c_void_p]
    st = L.accfft_amd_plan_dft_3d_r2c_slab(i3(8, 8, 10007), 2, None, None, af.F64, 0,
                                           ctypes.byref(h))
    assert st == 3
    # NULL plan
    assert L.accfft_amd_execute_r2c(None, None, None, None, 7) == 7
    assert L.accfft_amd_destroy_plan(None) == 7
    assert L.accfft_amd_grad(None, None, None, None, None, 7, None, 0) == 7
    # bad grid for the size query
    with pytest.raises(af.AccfftError):
        af.local_size_dft_r2c((8, 8, 8), af.F64, (2, 2), (2, 0))
    assert L.accfft_amd_status_string(3) == b"unsupported transform length"
    # slab share: the split must divide n0 and n1, n1/P a power of two, out of place
    slab = L.accfft_amd_plan_dft_3d_r2c_slab
    a, b = ctypes.c_void_p(16), ctypes.c_void_p(32)
    assert slab(i3(64, 64, 64), 0, a, b, af.F64, 0, ctypes.byref(h)) == 1
    assert slab(i3(60, 64, 64), 8, a, b, af.F64, 0, ctypes.byref(h)) == 3
    assert slab(i3(64, 60, 64), 8, a, b, af.F64, 0, ctypes.byref(h)) == 3
    assert slab(i3(48, 48, 64), 4, a, b, af.F64, 0, ctypes.byref(h)) == 3  # n1/P = 12
    assert slab(i3(64, 64, 64), 8, a, a, af.F64, 0, ctypes.byref(h)) == 3  # in place
    assert slab(i3(64, 64, 64), 8, None, None, af.F64, 0, ctypes.byref(h)) == 3
    assert not h.value
    assert L.accfft_amd_plan_set_exchange(None, af.EXCHANGE_FN(), None) == 7


def test_transpose_argument_checks(af):
    L = af.lib()
    # r == 0 or c == 0 is a no-op (src/transpose.cpp:667-668), negative sizes are invalid
    assert L.accfft_amd_local_transpose(0, 4, 8, None, None, 4, None) == 0
    assert L.accfft_amd_local_transpose(-1, 4, 8, None, None, 4, None) == 1
    assert L.accfft_amd_local_transpose_col(2, 2, 8, 8, None, None, 0, None) == 1
