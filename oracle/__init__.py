"""CPU FP64 oracle for the AccFFT single-rank hot path.

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` as the checker.  The product path in
``accfft_amd`` never imports this module.

Thin numpy/ctypes wrapper over ``oracle/oracle.c`` (built into
``oracle/liboracle.so`` by ``oracle/Makefile``).  Every function names the
reference code it restates; the restatement itself is documented in
``oracle.c``.

Arrays follow the reference's layouts (row-major, x slowest, z fastest):
real ``[N0][N1][N2]`` and half spectrum ``[N0][N1][N2//2+1]`` complex.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

(GRAD_X, GRAD_Y, GRAD_Z, LAPLACE, BIHARMONIC, INV_LAPLACE, INV_BIHARMONIC, POISSON_DISCRETE,
 POISSON_CONTINUOUS) = range(9)

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build(force: bool = False) -> str:
    """Compile oracle.c with gcc via oracle/Makefile."""
    if force or not os.path.exists(_LIB_PATH) or (
        os.path.getmtime(_LIB_PATH) < os.path.getmtime(os.path.join(_HERE, "oracle.c"))
    ):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_get_threads.restype = ctypes.c_int
        for name in ("orc_dft1d", "orc_dft1d_naive"):
            getattr(L, name).argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_int]
        L.orc_r2c_3d.argtypes = [_ip, _dp, _dp, ctypes.c_int]
        L.orc_c2r_3d.argtypes = [_ip, _dp, _dp, ctypes.c_int]
        L.orc_c2c_3d.argtypes = [_ip, ctypes.c_int, _dp, _dp, ctypes.c_int]
        L.orc_r2c_axis.argtypes = [_ip, ctypes.c_int, _dp, _dp]
        L.orc_c2r_axis.argtypes = [_ip, ctypes.c_int, _dp, _dp]
        L.orc_mult.argtypes = [ctypes.c_int, _ip, _ip, _ip, ctypes.c_double, _dp, _dp]
        L.orc_grad.argtypes = [_ip, _dp, _dp, _dp, _dp, ctypes.c_int]
        L.orc_grad_slow.argtypes = [_ip, _dp, _dp, _dp, _dp, ctypes.c_int]
        L.orc_spectral_op.argtypes = [ctypes.c_int, _ip, _dp, _dp]
        L.orc_divergence.argtypes = [_ip, _dp, _dp, _dp, _dp]
        L.orc_divergence_slow.argtypes = [_ip, _dp, _dp, _dp, _dp]
        for name in ("orc_local_transpose",):
            getattr(L, name).argtypes = [ctypes.c_int] * 3 + [ctypes.c_void_p] * 2 + [ctypes.c_int]
        for name in ("orc_local_transpose_ragged", "orc_local_transpose_col"):
            getattr(L, name).argtypes = [ctypes.c_int] * 4 + [ctypes.c_void_p] * 2 + [ctypes.c_int]
        for name in ("orc_local_size_r2c", "orc_local_size_c2c"):
            fn = getattr(L, name)
            fn.argtypes = [_ip] * 5 + [ctypes.c_int] * 5
            fn.restype = ctypes.c_int64
        _lib = L
    return _lib


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def get_threads() -> int:
    return lib().orc_get_threads()


def _d(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _n3(n) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(n, dtype=np.int32).reshape(3))


def _c(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.complex128))


def _r(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


# ---------------------------------------------------------------- 1-D
def dft1d(x, sign: int = -1) -> np.ndarray:
    """Mixed-radix Stockham DFT (unnormalised)."""
    x = _c(x).ravel()
    out = np.empty_like(x)
    lib().orc_dft1d(x.size, _d(x.view(np.float64)), _d(out.view(np.float64)), sign)
    return out


def dft1d_naive(x, sign: int = -1) -> np.ndarray:
    """O(n^2) DFT in long double (independent check of ``dft1d``)."""
    x = _c(x).ravel()
    out = np.empty_like(x)
    lib().orc_dft1d_naive(x.size, _d(x.view(np.float64)), _d(out.view(np.float64)), sign)
    return out


# ---------------------------------------------------------------- 3-D
def r2c(a, xyz: int = 7) -> np.ndarray:
    """3-D r2c, src/accfft_gpu.cpp:471-539 (z r2c, y c2c, x c2c)."""
    a = _r(a)
    n = _n3(a.shape)
    out = np.empty((n[0], n[1], n[2] // 2 + 1), np.complex128)
    lib().orc_r2c_3d(n.ctypes.data_as(_ip), _d(a), _d(out.view(np.float64)), xyz)
    return out


def c2r(h, n, xyz: int = 7) -> np.ndarray:
    """3-D c2r, src/accfft_gpu.cpp:540-603 (x c2c, y c2c, z c2r)."""
    h = _c(h)
    n = _n3(n)
    if xyz & 4:
        out = np.empty(tuple(n), np.float64)
    else:
        out = np.empty_like(h)
    lib().orc_c2r_3d(n.ctypes.data_as(_ip), _d(h.view(np.float64)), _d(out.view(np.float64)), xyz)
    return out


def c2c(a, direction: int, xyz: int = 7) -> np.ndarray:
    """3-D c2c, src/accfft_gpu.cpp:943-1136."""
    a = _c(a)
    n = _n3(a.shape)
    out = np.empty_like(a)
    lib().orc_c2c_3d(n.ctypes.data_as(_ip), int(direction), _d(a.view(np.float64)),
                     _d(out.view(np.float64)), xyz)
    return out


def r2c_axis(a, axis: int) -> np.ndarray:
    """Per-axis r2c, src/accfft_gpu.cpp:1212-1604."""
    a = _r(a)
    n = _n3(a.shape)
    shape = list(n)
    shape[axis] = n[axis] // 2 + 1
    out = np.empty(shape, np.complex128)
    lib().orc_r2c_axis(n.ctypes.data_as(_ip), axis, _d(a), _d(out.view(np.float64)))
    return out


def c2r_axis(h, n, axis: int) -> np.ndarray:
    h = _c(h)
    n = _n3(n)
    out = np.empty(tuple(n), np.float64)
    lib().orc_c2r_axis(n.ctypes.data_as(_ip), axis, _d(h.view(np.float64)), _d(out))
    return out


# ---------------------------------------------------------------- operators
def mult(kind: int, N, h, ostart=(0, 0, 0), scale: float = 1.0) -> np.ndarray:
    """Spectral multiplier over block ``h`` at global offset ``ostart``
    (src/operators_cuda.cu:15-205, src/operators.txx:276-392)."""
    h = _c(h)
    N = _n3(N)
    osize = _n3(h.shape)
    ostart = _n3(ostart)
    out = np.empty_like(h)
    lib().orc_mult(kind, N.ctypes.data_as(_ip), osize.ctypes.data_as(_ip),
                   ostart.ctypes.data_as(_ip), float(scale), _d(h.view(np.float64)),
                   _d(out.view(np.float64)))
    return out


def grad(a, mask: int = 7, slow: bool = False):
    """accfft_grad_gpu_t (fast, per axis; operators_gpu.txx:172-278) or the
    3-D ``_slow`` variant (:72-170).  Returns (gx, gy, gz), None where the
    mask bit is clear."""
    a = _r(a)
    n = _n3(a.shape)
    outs = [np.empty_like(a) if mask & (1 << d) else None for d in range(3)]
    ptrs = [_d(o) if o is not None else None for o in outs]
    fn = lib().orc_grad_slow if slow else lib().orc_grad
    fn(n.ctypes.data_as(_ip), _d(a), ptrs[0], ptrs[1], ptrs[2], mask)
    return tuple(outs)


def spectral_op(kind: int, a) -> np.ndarray:
    a = _r(a)
    n = _n3(a.shape)
    out = np.empty_like(a)
    lib().orc_spectral_op(kind, n.ctypes.data_as(_ip), _d(a), _d(out))
    return out


def laplace(a):
    """accfft_laplace_gpu_t, operators_gpu.txx:280-334."""
    return spectral_op(LAPLACE, a)


def biharmonic(a):
    """accfft_biharmonic_gpu_t, operators_gpu.txx:555-609."""
    return spectral_op(BIHARMONIC, a)


def inv_laplace(a):
    """accfft_inv_laplace_t, operators.txx:1270-1321 (CPU-only upstream)."""
    return spectral_op(INV_LAPLACE, a)


def inv_biharmonic(a):
    """accfft_inv_biharmonic_t multiplier, operators.txx:336-392."""
    return spectral_op(INV_BIHARMONIC, a)


def poisson_solve(rho, method: int = 0):
    """The step4 FFT Poisson solve, steps/step4/step4_gpu.cpp:520-640: r2c,
    poisson_fourier_filter (steps/step4/kernels.cu:24-110), c2r."""
    return spectral_op(POISSON_DISCRETE + int(method), rho)


def poisson_filter(h, N, ostart=(0, 0, 0), method: int = 0) -> np.ndarray:
    """poisson_fourier_filter_gpu on a half spectrum block at ``ostart``
    (steps/step4/kernels.cu:24-146); returns the filtered copy."""
    N = _n3(N)
    return mult(POISSON_DISCRETE + int(method), N, h, ostart, 1.0 / float(np.prod(N)))


def divergence(ax, ay, az, slow: bool = False):
    """accfft_divergence_gpu_t (:448-553) or ``_slow_t`` (:336-446)."""
    ax, ay, az = _r(ax), _r(ay), _r(az)
    n = _n3(ax.shape)
    out = np.empty_like(ax)
    fn = lib().orc_divergence_slow if slow else lib().orc_divergence
    fn(n.ctypes.data_as(_ip), _d(ax), _d(ay), _d(az), _d(out))
    return out


# ---------------------------------------------------------------- slab share
def slab_r2c(a, P: int) -> np.ndarray:
    """One rank's share of the r2c on a P x 1 grid with the exchange taken as
    the identity (src/accfft_gpu.cpp:471-539; T_plan_2 pack of
    src/transpose_gpu.cpp:3889-3891): z and y transforms of the x-slab
    ``a`` [n0][N1][N2], the pack local_transpose_col(n0, P, n1*Nh), then the
    x transform of the received [P*n0][n1][Nh]."""
    a = _r(a)
    n0, N1, N2 = a.shape
    n1, Nh = N1 // P, N2 // 2 + 1
    h = r2c(a, xyz=6)
    packed = local_transpose_col(n0, P, n1 * Nh, n1 * Nh, h.ravel()).reshape(P * n0, n1, Nh)
    return c2c(packed, -1, xyz=1)


def slab_c2r(h, N2: int, P: int) -> np.ndarray:
    """The inverse: x transform of [N0][n1][Nh], the unpack
    local_transpose(P, n0, n1*Nh) (src/transpose_gpu.cpp:1411-1413), y and z
    c2r of the x-slab [n0][N1][N2]."""
    h = _c(h)
    N0, n1, Nh = h.shape
    n0 = N0 // P
    x = c2c(h, +1, xyz=1)
    un = local_transpose(P, n0, n1 * Nh, x.ravel()).reshape(n0, P * n1, Nh)
    return c2r(un, (n0, P * n1, N2), xyz=6)


# ---------------------------------------------------------------- adjoints
# No counterpart in the reference: adjoints of its real-linear maps for
# reverse-mode differentiation, with <a, b> = Re sum conj(a) b on half
# spectra.  Pinned in tests/test_oracle.py against explicit transposed
# matrices built column by column from r2c / c2r / grad / divergence.
def adjoint_weight(n) -> np.ndarray:
    """w[k] over the z half spectrum: 1 at DC and (even n[2]) Nyquist, 1/2 elsewhere."""
    n2 = int(n[2])
    w = np.full(n2 // 2 + 1, 0.5)
    w[0] = 1.0
    if n2 % 2 == 0:
        w[-1] = 1.0
    return w


def adjoint_r2c(G, n) -> np.ndarray:
    """r2c^T(G) = c2r(w G)."""
    return c2r(_c(G) * adjoint_weight(n), n)


def adjoint_c2r(y) -> np.ndarray:
    """c2r^T(y) = r2c(y) / w."""
    y = _r(y)
    return r2c(y) / adjoint_weight(y.shape)


def adjoint_grad(gx, gy, gz) -> np.ndarray:
    """grad^T(gx, gy, gz) = -div(gx, gy, gz) (per-axis derivatives)."""
    return -divergence(gx, gy, gz)


def adjoint_divergence(v):
    """div^T(v) = -grad(v) (per-axis derivatives)."""
    return tuple(-g for g in grad(v))


# ---------------------------------------------------------------- transposes
def local_transpose(r, c, t, a) -> np.ndarray:
    """[r][c][t] -> [c][r][t]; src/transpose.cpp:664-679 (cuda1/cuda3)."""
    a = np.ascontiguousarray(a)
    out = np.empty(r * c * t, a.dtype)
    lib().orc_local_transpose(r, c, t, a.ctypes.data, out.ctypes.data, a.dtype.itemsize)
    return out


def local_transpose_ragged(r, c, t, t2, a) -> np.ndarray:
    """Unpack with ragged last row; src/transpose.cpp:681-703 (cuda2)."""
    a = np.ascontiguousarray(a)
    out = np.empty(c * ((r - 1) * t + t2), a.dtype)
    lib().orc_local_transpose_ragged(r, c, t, t2, a.ctypes.data, out.ctypes.data,
                                     a.dtype.itemsize)
    return out


def local_transpose_col(r, c, t, t2, a) -> np.ndarray:
    """Pack with ragged last column; src/transpose.cpp:705-729 (cuda4)."""
    a = np.ascontiguousarray(a)
    out = np.empty(r * ((c - 1) * t + t2), a.dtype)
    lib().orc_local_transpose_col(r, c, t, t2, a.ctypes.data, out.ctypes.data,
                                  a.dtype.itemsize)
    return out


# ---------------------------------------------------------------- sizes
def pencil_pack(nprocs_1, howmany, local_n0, n_tuples, local_n1_proc, data, idist, N1,
                local_1_start_proc) -> np.ndarray:
    """memcpy_v1_h1, src/transpose_cuda.cu:97-143: segments (proc, h, i),
    each the local_n1_proc[proc] tuples of row i of batch h starting at
    column local_1_start_proc[proc], concatenated."""
    data = np.asarray(data).ravel()
    out = []
    for p in range(nprocs_1):
        for h in range(howmany):
            for i in range(local_n0):
                src = h * idist + (i * N1 + local_1_start_proc[p]) * n_tuples
                out.append(data[src:src + n_tuples * local_n1_proc[p]])
    return np.concatenate(out) if out else np.zeros(0, data.dtype)


def pencil_unpack(nprocs_0, howmany, local_0_start_proc, local_n0_proc, data, odist, local_n1,
                  n_tuples, send_recv) -> np.ndarray:
    """memcpy_v1_h2, src/transpose_cuda.cu:145-197: consecutive segments of
    local_n1 * n_tuples elements of ``send_recv``, (proc, h, i) for i in
    proc's row range, written to data[h * odist + i * local_n1 * n_tuples];
    returns the updated copy of ``data``."""
    data = np.array(data, copy=True).ravel()
    send_recv = np.asarray(send_recv).ravel()
    seg = local_n1 * n_tuples
    ptr = 0
    for p in range(nprocs_0):
        for h in range(howmany):
            for i in range(local_0_start_proc[p], local_0_start_proc[p] + local_n0_proc[p]):
                dst = h * odist + i * seg
                data[dst:dst + seg] = send_recv[ptr:ptr + seg]
                ptr += seg
    return data


def local_size_r2c(n, elem: int, p=(1, 1), coords=(0, 0)):
    """accfft_local_size_dft_r2c_t<T>, src/accfft_common.cpp:98-192.
    Returns (alloc_max_bytes, isize, istart, osize, ostart)."""
    n = _n3(n)
    arrs = [np.zeros(3, np.int32) for _ in range(4)]
    amax = lib().orc_local_size_r2c(n.ctypes.data_as(_ip), *[x.ctypes.data_as(_ip) for x in arrs],
                                    p[0], p[1], coords[0], coords[1], elem)
    return (int(amax),) + tuple(tuple(int(v) for v in x) for x in arrs)


def local_size_c2c(n, elem: int, p=(1, 1), coords=(0, 0)):
    """accfft_local_size_dft_c2c_t<T>, src/accfft_common.cpp:203-256."""
    n = _n3(n)
    arrs = [np.zeros(3, np.int32) for _ in range(4)]
    amax = lib().orc_local_size_c2c(n.ctypes.data_as(_ip), *[x.ctypes.data_as(_ip) for x in arrs],
                                    p[0], p[1], coords[0], coords[1], elem)
    return (int(amax),) + tuple(tuple(int(v) for v in x) for x in arrs)
