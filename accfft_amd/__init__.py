"""accfft_amd -- MI355X (gfx950) implementation of AccFFT's single-GPU hot path.

Python mirror of the reference's GPU C++ API (include/accfft_gpu.h:97-155,
include/accfft_gpuf.h:97-158, include/accfft_operators_gpu.h:31-67) over the
C ABI of ``libaccfft_amd.so`` (declared in include/accfft_amd.h).  Device
memory is handled with torch tensors; every call is ordered on torch's
current HIP stream.

The native library is mandatory: if ``libaccfft_amd.so`` is missing or a GPU
call fails, these functions raise -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libaccfft_amd.so")

ACCFFT_FORWARD = -1
ACCFFT_BACKWARD = +1
ACCFFT_INVERSE = +1
ACCFFT_ESTIMATE = 1
ACCFFT_MEASURE = 2
ACCFFT_PATIENT = 4
BLOCKING = 0x100
PRESERVE_INPUT = 0x200

F64, F32 = 0, 1
OK = 0

(GRAD_X, GRAD_Y, GRAD_Z, LAPLACE, BIHARMONIC, INV_LAPLACE, INV_BIHARMONIC, POISSON_DISCRETE,
 POISSON_CONTINUOUS) = range(9)

_lib = None
_i3 = ctypes.c_int * 3
_vp = ctypes.c_void_p


ERR_ALLOC = 6  # ACCFFT_AMD_ERR_ALLOC: a device allocation failed


class AccfftError(RuntimeError):
    """A failed call; ``status`` is the C ABI's status code (None for
    argument checks made before the call)."""

    def __init__(self, msg: str, status: Optional[int] = None):
        super().__init__(msg)
        self.status = status


# int (*)(const void* send, void* recv, int64_t block_bytes, int nprocs,
#         int direction, void* stream, void* user)
EXCHANGE_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                               ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p)

# int (*)(int comm, const void* send, const int64_t* send_bytes,
#         const int64_t* send_offsets, void* recv, const int64_t* recv_bytes,
#         const int64_t* recv_offsets, int nprocs, int direction, void* stream,
#         void* user)
_I64P = ctypes.POINTER(ctypes.c_int64)
EXCHANGE_V_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_int, ctypes.c_void_p, _I64P, _I64P,
                                 ctypes.c_void_p, _I64P, _I64P, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_void_p, ctypes.c_void_p)


def lib() -> ctypes.CDLL:
    """Loads libaccfft_amd.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise AccfftError(
            f"{LIB_PATH} is missing: build it with `python -m accfft_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    i64p = ctypes.POINTER(ctypes.c_int64)
    dp = ctypes.POINTER(ctypes.c_double)
    L.accfft_amd_status_string.restype = ctypes.c_char_p
    L.accfft_amd_status_string.argtypes = [ctypes.c_int]
    for nm in ("accfft_amd_local_size_dft_r2c", "accfft_amd_local_size_dft_c2c"):
        getattr(L, nm).argtypes = [_i3] * 5 + [ctypes.c_int] * 5 + [i64p]
    for nm in ("accfft_amd_plan_dft_3d_r2c", "accfft_amd_plan_dft_3d_c2c"):
        getattr(L, nm).argtypes = [_i3, _vp, _vp, ctypes.c_int, ctypes.c_uint, ctypes.POINTER(_vp)]
    L.accfft_amd_plan_dft_3d_r2c_slab.argtypes = [_i3, ctypes.c_int, _vp, _vp, ctypes.c_int,
                                                  ctypes.c_uint, ctypes.POINTER(_vp)]
    L.accfft_amd_plan_set_exchange.argtypes = [_vp, EXCHANGE_FN, _vp]
    L.accfft_amd_plan_dft_3d_r2c_pencil.argtypes = [_i3] + [ctypes.c_int] * 4 + [
        _vp, _vp, ctypes.c_int, ctypes.c_uint, ctypes.POINTER(_vp)]
    L.accfft_amd_plan_set_exchange_v.argtypes = [_vp, EXCHANGE_V_FN, _vp]
    L.accfft_amd_destroy_plan.argtypes = [_vp]
    L.accfft_amd_plan_set_stream.argtypes = [_vp, _vp]
    L.accfft_amd_plan_sync.argtypes = [_vp]
    L.accfft_amd_plan_info.argtypes = [_vp, _i3, _i3, _i3, _i3, _i3, i64p,
                                       ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                       ctypes.POINTER(ctypes.c_int)]
    L.accfft_amd_execute_r2c.argtypes = [_vp, _vp, _vp, dp, ctypes.c_uint]
    L.accfft_amd_execute_c2r.argtypes = [_vp, _vp, _vp, dp, ctypes.c_uint]
    L.accfft_amd_execute_c2c.argtypes = [_vp, ctypes.c_int, _vp, _vp, dp, ctypes.c_uint]
    L.accfft_amd_execute_r2c_axis.argtypes = [_vp, ctypes.c_int, _vp, _vp, dp]
    L.accfft_amd_execute_c2r_axis.argtypes = [_vp, ctypes.c_int, _vp, _vp, dp]
    L.accfft_amd_grad.argtypes = [_vp, _vp, _vp, _vp, _vp, ctypes.c_uint, dp, ctypes.c_int]
    for nm in ("accfft_amd_laplace", "accfft_amd_biharmonic", "accfft_amd_inv_laplace",
               "accfft_amd_inv_biharmonic"):
        getattr(L, nm).argtypes = [_vp, _vp, _vp, dp]
    L.accfft_amd_divergence.argtypes = [_vp, _vp, _vp, _vp, _vp, dp, ctypes.c_int]
    L.accfft_amd_poisson_solve.argtypes = [_vp, ctypes.c_int, _vp, _vp, dp]
    L.accfft_amd_poisson_filter.argtypes = [ctypes.c_int, _vp, _i3, _i3, _i3, ctypes.c_int, _vp]
    L.accfft_amd_mult_wave_number.argtypes = [ctypes.c_int, ctypes.c_int, _vp, _vp, _i3, _i3,
                                              _i3, ctypes.c_uint, _vp]
    L.accfft_amd_axpy.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_double, _vp, _vp, _vp]
    L.accfft_amd_local_transpose.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int64, _vp,
                                             _vp, ctypes.c_int, _vp]
    for nm in ("accfft_amd_local_transpose_ragged", "accfft_amd_local_transpose_col"):
        getattr(L, nm).argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.c_int64,
                                   _vp, _vp, ctypes.c_int, _vp]
    L.accfft_amd_local_transpose_inplace.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int64,
                                                     _vp, ctypes.c_int, _vp]
    L.accfft_amd_pencil_pack.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int64,
                                         i64p, _vp, _vp, ctypes.c_int64, ctypes.c_int64, i64p,
                                         ctypes.c_int, _vp]
    L.accfft_amd_pencil_unpack.argtypes = [ctypes.c_int, ctypes.c_int, i64p, i64p, _vp,
                                           ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, _vp,
                                           ctypes.c_int, _vp]
    L.accfft_amd_adjoint_r2c.argtypes = [_vp, _vp, _vp, dp]
    L.accfft_amd_adjoint_c2r.argtypes = [_vp, _vp, _vp, dp]
    L.accfft_amd_adjoint_grad.argtypes = [_vp, _vp, _vp, _vp, _vp, dp]
    L.accfft_amd_adjoint_divergence.argtypes = [_vp, _vp, _vp, _vp, _vp, dp]
    L.accfft_amd_plan_release_workspace.argtypes = [_vp]
    L.accfft_amd_plan_get_tuning.argtypes = [_vp, _i3]
    L.accfft_amd_plan_set_tuning.argtypes = [_vp, ctypes.c_int, ctypes.c_int]
    L.accfft_amd_plan_workspace_bytes.argtypes = [_vp, i64p]
    L.accfft_amd_last_elapsed_ms.argtypes = [_vp, ctypes.POINTER(ctypes.c_float)]
    L.accfft_amd_plan_set_profiling.argtypes = [_vp, ctypes.c_int]
    L.accfft_amd_profile_reset.argtypes = [_vp]
    L.accfft_amd_profile_count.argtypes = [_vp, ctypes.POINTER(ctypes.c_int)]
    L.accfft_amd_profile_get.argtypes = [_vp, ctypes.c_int, ctypes.c_char_p, ctypes.c_int,
                                         dp, i64p, dp]
    _lib = L
    return L


def _check(status: int, what: str) -> None:
    if status != OK:
        msg = lib().accfft_amd_status_string(status).decode()
        raise AccfftError(f"{what}: {msg} (status {status})", status)


def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t) -> Optional[int]:
    if t is None:
        return None
    return t.data_ptr()


def _family(precision: int):
    import torch
    return ((torch.float32, torch.complex64) if precision == F32
            else (torch.float64, torch.complex128))


def _arg(t, nbytes: int, precision: int, what: str, device: Optional[int] = None,
         optional: bool = True) -> Optional[int]:
    """Device pointer of ``t`` after checking that the kernels may use it:
    a contiguous GPU tensor (on ``device`` when given) of the precision's
    real or complex dtype holding at least ``nbytes`` bytes.  ``None`` passes
    through where the C ABI takes NULL (the plan's own buffers)."""
    import torch
    if t is None:
        if optional:
            return None
        raise AccfftError(f"{what}: a tensor is required")
    if not isinstance(t, torch.Tensor):
        raise AccfftError(f"{what}: expected a torch.Tensor, got {type(t).__name__}")
    if not t.is_cuda:
        raise AccfftError(f"{what}: tensor is on {t.device}, not a GPU")
    if device is not None and t.device.index != device:
        raise AccfftError(f"{what}: tensor is on cuda:{t.device.index}, the plan on cuda:{device}")
    if not t.is_contiguous():
        raise AccfftError(f"{what}: tensor is not contiguous (pass .contiguous())")
    if precision is not None and t.dtype not in _family(precision):
        raise AccfftError(f"{what}: dtype {t.dtype} does not match the plan's precision "
                          f"({'FP32' if precision == F32 else 'FP64'})")
    have = t.numel() * t.element_size()
    if have < nbytes:
        raise AccfftError(f"{what}: tensor holds {have} bytes, the call needs {nbytes}")
    return t.data_ptr()


def _gpu(t, what: str, nbytes: int = 0):
    """Checks an argument of the element-wise and copy kernels (any element
    type): a contiguous GPU tensor of at least ``nbytes`` bytes."""
    return _arg(t, nbytes, None, what, None, False)


def _fp_of(dtype) -> int:
    import torch
    if dtype in (torch.float32, torch.complex64):
        return F32
    if dtype in (torch.float64, torch.complex128):
        return F64
    raise AccfftError(f"unsupported dtype {dtype}")


# ------------------------------------------------------------------ sizes
def local_size_dft_r2c(n: Sequence[int], precision: int = F64, grid=(1, 1), coords=(0, 0)):
    """accfft_local_size_dft_r2c_gpu(f) (src/accfft_gpu.cpp:62-65).
    Returns (alloc_max_bytes, isize, istart, osize, ostart)."""
    arr = [_i3() for _ in range(4)]
    amax = ctypes.c_int64()
    _check(lib().accfft_amd_local_size_dft_r2c(_i3(*n), *arr, precision, grid[0], grid[1],
                                                 coords[0], coords[1], ctypes.byref(amax)),
           "local_size_dft_r2c")
    return (amax.value,) + tuple(tuple(a) for a in arr)


def local_size_dft_c2c(n: Sequence[int], precision: int = F64, grid=(1, 1), coords=(0, 0)):
    """accfft_local_size_dft_c2c_gpu(f) (src/accfft_gpu.cpp:689-692)."""
    arr = [_i3() for _ in range(4)]
    amax = ctypes.c_int64()
    _check(lib().accfft_amd_local_size_dft_c2c(_i3(*n), *arr, precision, grid[0], grid[1],
                                                 coords[0], coords[1], ctypes.byref(amax)),
           "local_size_dft_c2c")
    return (amax.value,) + tuple(tuple(a) for a in arr)


# ------------------------------------------------------------------ plan
class Plan:
    """accfft_plan_gpu(f) on a 1x1 grid (include/accfft_gpu.h:37-95).

    As upstream (src/accfft_gpu.cpp:99-103) the plan is in place -- real rows
    padded to 2*(n[2]//2+1) -- exactly when ``data`` and ``data_out`` are the
    same buffer, which includes passing neither."""

    def __init__(self, kind: str, n: Sequence[int], data=None, data_out=None,
                 precision: int = F64, flags: int = ACCFFT_MEASURE, nprocs: int = 1,
                 grid=(1, 1), coords=(0, 0)):
        """``kind`` is "r2c", "c2c", "r2c_slab" (one rank's share of an r2c
        on an ``nprocs`` x 1 grid, ``n`` global) or "r2c_pencil" (the share
        of the rank at ``coords`` of a ``grid`` = (p0, p1) process grid,
        the reference's accfft_plan_dft_3d_r2c_gpu on c_dims = grid); see
        include/accfft_amd.h."""
        self.kind = kind
        self.n = tuple(int(v) for v in n)
        self.precision = precision
        self.nprocs = nprocs
        self.grid, self.coords = tuple(grid), tuple(coords)
        h = _vp()
        if kind == "r2c_slab":
            _check(lib().accfft_amd_plan_dft_3d_r2c_slab(_i3(*self.n), nprocs, _ptr(data),
                                                         _ptr(data_out), precision, flags,
                                                         ctypes.byref(h)), "plan_dft_3d_r2c_slab")
        elif kind == "r2c_pencil":
            _check(lib().accfft_amd_plan_dft_3d_r2c_pencil(
                _i3(*self.n), self.grid[0], self.grid[1], self.coords[0], self.coords[1],
                _ptr(data), _ptr(data_out), precision, flags, ctypes.byref(h)),
                "plan_dft_3d_r2c_pencil")
        else:
            fn = (lib().accfft_amd_plan_dft_3d_r2c if kind == "r2c"
                  else lib().accfft_amd_plan_dft_3d_c2c)
            _check(fn(_i3(*self.n), _ptr(data), _ptr(data_out), precision, flags,
                      ctypes.byref(h)), f"plan_dft_3d_{kind}")
        self._exchange = None
        self._h = h
        isz, ist, osz, ost = _i3(), _i3(), _i3(), _i3()
        amax = ctypes.c_int64()
        fp, knd, inp = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(lib().accfft_amd_plan_info(h, _i3(), isz, ist, osz, ost, ctypes.byref(amax),
                                          ctypes.byref(fp), ctypes.byref(knd), ctypes.byref(inp)),
               "plan_info")
        self.isize, self.istart = tuple(isz), tuple(ist)
        self.osize, self.ostart = tuple(osz), tuple(ost)
        self.alloc_max = amax.value
        self.inplace = bool(inp.value)
        self._keep = (data, data_out)
        import torch
        self.device = torch.cuda.current_device() if torch.cuda.is_available() else None
        s = 4 if precision == F32 else 8
        N0, N1, N2 = self.n
        nh = N2 // 2 + 1
        if kind == "c2c":
            self._in_bytes = N0 * N1 * N2 * 2 * s
        else:
            z = 2 * nh if self.inplace else self.isize[2]
            self._in_bytes = self.isize[0] * self.isize[1] * z * s
        self._out_bytes = self.osize[0] * self.osize[1] * self.osize[2] * 2 * s

    def _a(self, t, nbytes, what, optional=True):
        return _arg(t, nbytes, self.precision, what, self.device, optional)

    def _field(self, t, what, optional=True):
        """A real field [N0][N1][N2] of the operators."""
        s = 4 if self.precision == F32 else 8
        return self._a(t, self.n[0] * self.n[1] * self.n[2] * s, what, optional)

    def _axis_bytes(self, axis: int) -> int:
        """Bytes of the half spectrum of a per-axis r2c along ``axis``."""
        s = 4 if self.precision == F32 else 8
        m = list(self.n)
        m[axis] = m[axis] // 2 + 1
        return m[0] * m[1] * m[2] * 2 * s

    @property
    def handle(self):
        return self._h

    def destroy(self):
        if self._h:
            lib().accfft_amd_destroy_plan(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def _bind(self):
        if not self._h:
            raise AccfftError("plan destroyed")
        _check(lib().accfft_amd_plan_set_stream(self._h, _stream()), "set_stream")
        return self._h

    @staticmethod
    def _timer(timer):
        if timer is None:
            return None, None
        arr = (ctypes.c_double * len(timer))(*timer)
        return arr, ctypes.cast(arr, ctypes.POINTER(ctypes.c_double))

    @staticmethod
    def _untimer(timer, arr):
        if timer is not None:
            for i in range(len(timer)):
                timer[i] = arr[i]

    def execute_r2c(self, data=None, data_out=None, timer=None, xyz: int = 7):
        arr, tp = self._timer(timer)
        out_b = self._out_bytes if xyz & 4 else 0
        _check(lib().accfft_amd_execute_r2c(self._bind(), self._a(data, self._in_bytes, "r2c input"),
                                            self._a(data_out, out_b, "r2c output"), tp, xyz),
               "execute_r2c")
        self._untimer(timer, arr)

    def execute_c2r(self, data=None, data_out=None, timer=None, xyz: int = 7):
        arr, tp = self._timer(timer)
        out_b = self._in_bytes if xyz & 4 else 0
        _check(lib().accfft_amd_execute_c2r(self._bind(), self._a(data, self._out_bytes, "c2r input"),
                                            self._a(data_out, out_b, "c2r output"), tp, xyz),
               "execute_c2r")
        self._untimer(timer, arr)

    def execute_c2c(self, direction: int, data=None, data_out=None, timer=None, xyz: int = 7):
        arr, tp = self._timer(timer)
        _check(lib().accfft_amd_execute_c2c(self._bind(), direction,
                                            self._a(data, self._in_bytes, "c2c input"),
                                            self._a(data_out, self._out_bytes, "c2c output"),
                                            tp, xyz), "execute_c2c")
        self._untimer(timer, arr)

    def execute_r2c_axis(self, axis: int, data, data_out, timer=None):
        arr, tp = self._timer(timer)
        _check(lib().accfft_amd_execute_r2c_axis(
            self._bind(), axis, self._field(data, "r2c_axis input", False),
            self._a(data_out, self._axis_bytes(axis), "r2c_axis output", False), tp),
            "execute_r2c_axis")
        self._untimer(timer, arr)

    def execute_c2r_axis(self, axis: int, data, data_out, timer=None):
        arr, tp = self._timer(timer)
        _check(lib().accfft_amd_execute_c2r_axis(
            self._bind(), axis, self._a(data, self._axis_bytes(axis), "c2r_axis input", False),
            self._field(data_out, "c2r_axis output", False), tp), "execute_c2r_axis")
        self._untimer(timer, arr)

    # operators (include/accfft_operators_gpu.h:44-67)
    def grad(self, A, Ax=None, Ay=None, Az=None, mask: int = 7, timer=None, slow=False):
        arr, tp = self._timer(timer)
        _check(lib().accfft_amd_grad(self._bind(), self._field(Ax, "grad x output"),
                                     self._field(Ay, "grad y output"),
                                     self._field(Az, "grad z output"),
                                     self._field(A, "grad input", False), mask, tp, int(slow)),
               "grad")
        self._untimer(timer, arr)

    def _spectral(self, name, out, A, timer):
        arr, tp = self._timer(timer)
        _check(getattr(lib(), name)(self._bind(), self._field(out, name + " output", False),
                                    self._field(A, name + " input", False), tp), name)
        self._untimer(timer, arr)

    def laplace(self, A, out, timer=None):
        self._spectral("accfft_amd_laplace", out, A, timer)

    def biharmonic(self, A, out, timer=None):
        self._spectral("accfft_amd_biharmonic", out, A, timer)

    def inv_laplace(self, A, out, timer=None):
        self._spectral("accfft_amd_inv_laplace", out, A, timer)

    def inv_biharmonic(self, A, out, timer=None):
        self._spectral("accfft_amd_inv_biharmonic", out, A, timer)

    def poisson_solve(self, rho, phi, method: int = 0, timer=None):
        """The step4 FFT Poisson solve (steps/step4/step4_gpu.cpp:520-640) with
        the Fourier filter (steps/step4/kernels.cu:24-110) fused into the x
        pass: phi = zero-mean solution of laplace(phi) = rho on the unit box."""
        arr, tp = self._timer(timer)
        _check(lib().accfft_amd_poisson_solve(self._bind(), int(method),
                                              self._field(phi, "poisson_solve output", False),
                                              self._field(rho, "poisson_solve input", False), tp),
               "poisson_solve")
        self._untimer(timer, arr)

    def divergence(self, Ax, Ay, Az, out, timer=None, slow=False):
        arr, tp = self._timer(timer)
        _check(lib().accfft_amd_divergence(self._bind(), self._field(out, "divergence output", False),
                                           self._field(Ax, "divergence x input"),
                                           self._field(Ay, "divergence y input"),
                                           self._field(Az, "divergence z input"), tp, int(slow)),
               "divergence")
        self._untimer(timer, arr)

    def set_exchange(self, fn):
        """Exchange hook of a slab plan: ``fn(send, recv, block_bytes, nprocs,
        direction, stream)`` (ints) returning a status; None = identity."""
        if fn is None:
            cb = EXCHANGE_FN()
        else:
            def tramp(send, recv, nbytes, nprocs, direction, stream, user):
                try:
                    return int(fn(send, recv, nbytes, nprocs, direction, stream))
                except Exception:  # a Python error must not unwind through C
                    return 5
            cb = EXCHANGE_FN(tramp)
        _check(lib().accfft_amd_plan_set_exchange(self._h, cb, None), "set_exchange")
        self._exchange = cb  # keep the thunk alive while the plan holds it

    def set_exchange_v(self, fn):
        """Exchange hook of a pencil plan: ``fn(comm, send, send_bytes,
        send_offsets, recv, recv_bytes, recv_offsets, direction, stream)``
        (pointers and the stream as ints, the per-peer counts and offsets as
        lists of ints; comm 0 = the row communicator of the p1 ranks sharing
        this rank's c0, 1 = the column communicator of the p0 ranks sharing
        its c1) returning a status; None = identity."""
        if fn is None:
            cb = EXCHANGE_V_FN()
        else:
            def tramp(comm, send, sb, so, recv, rb, ro, nprocs, direction, stream, user):
                try:
                    lst = lambda a: [a[i] for i in range(nprocs)]
                    return int(fn(comm, send, lst(sb), lst(so), recv, lst(rb), lst(ro),
                                  direction, stream))
                except Exception:  # a Python error must not unwind through C
                    return 5
            cb = EXCHANGE_V_FN(tramp)
        _check(lib().accfft_amd_plan_set_exchange_v(self._h, cb, None), "set_exchange_v")
        self._exchange = cb  # keep the thunk alive while the plan holds it

    # adjoints of the real-linear maps above (no counterpart in the reference)
    def adjoint_r2c(self, G, x, timer=None):
        """x = r2c^T(G): the gradient of <G, r2c(x)> with respect to x."""
        arr, tp = self._timer(timer)
        _check(lib().accfft_amd_adjoint_r2c(self._bind(),
                                            self._a(G, self._out_bytes, "adjoint_r2c input", False),
                                            self._field(x, "adjoint_r2c output", False), tp),
               "adjoint_r2c")
        self._untimer(timer, arr)

    def adjoint_c2r(self, y, G, timer=None):
        """G = c2r^T(y); ``y`` is not overwritten."""
        arr, tp = self._timer(timer)
        _check(lib().accfft_amd_adjoint_c2r(self._bind(), self._field(y, "adjoint_c2r input", False),
                                            self._a(G, self._out_bytes, "adjoint_c2r output", False),
                                            tp), "adjoint_c2r")
        self._untimer(timer, arr)

    def adjoint_grad(self, gx, gy, gz, out, timer=None):
        """out = grad^T(gx, gy, gz) = -div(gx, gy, gz); None components are skipped."""
        arr, tp = self._timer(timer)
        _check(lib().accfft_amd_adjoint_grad(self._bind(), self._field(out, "adjoint_grad output", False),
                                             self._field(gx, "adjoint_grad x input"),
                                             self._field(gy, "adjoint_grad y input"),
                                             self._field(gz, "adjoint_grad z input"), tp),
               "adjoint_grad")
        self._untimer(timer, arr)

    def adjoint_divergence(self, v, ox=None, oy=None, oz=None, timer=None):
        """(ox, oy, oz) = div^T(v) = -grad(v); None outputs are skipped."""
        arr, tp = self._timer(timer)
        _check(lib().accfft_amd_adjoint_divergence(
            self._bind(), self._field(ox, "adjoint_divergence x output"),
            self._field(oy, "adjoint_divergence y output"),
            self._field(oz, "adjoint_divergence z output"),
            self._field(v, "adjoint_divergence input", False), tp), "adjoint_divergence")
        self._untimer(timer, arr)

    def tuning(self):
        """Tile configuration per axis (-1 rule, -2 one-level, v >= 0 variant)."""
        c = _i3()
        _check(lib().accfft_amd_plan_get_tuning(self._bind(), c), "plan_get_tuning")
        return tuple(c)

    def set_tuning(self, axis: int, choice: int):
        _check(lib().accfft_amd_plan_set_tuning(self._bind(), axis, choice), "plan_set_tuning")

    def release_workspace(self):
        _check(lib().accfft_amd_plan_release_workspace(self._bind()), "release_workspace")

    def workspace_bytes(self) -> int:
        v = ctypes.c_int64()
        _check(lib().accfft_amd_plan_workspace_bytes(self._h, ctypes.byref(v)), "workspace_bytes")
        return v.value

    def sync(self):
        _check(lib().accfft_amd_plan_sync(self._bind()), "sync")

    def set_profiling(self, enable: bool = True):
        _check(lib().accfft_amd_plan_set_profiling(self._h, int(enable)), "set_profiling")

    def profile_reset(self):
        _check(lib().accfft_amd_profile_reset(self._h), "profile_reset")

    def profile(self) -> dict:
        """{kernel kind: (total_ms, launches, algorithmic bytes per launch)}."""
        n = ctypes.c_int()
        _check(lib().accfft_amd_profile_count(self._h, ctypes.byref(n)), "profile_count")
        out = {}
        for i in range(n.value):
            buf = ctypes.create_string_buffer(128)
            ms, cnt, b = ctypes.c_double(), ctypes.c_int64(), ctypes.c_double()
            _check(lib().accfft_amd_profile_get(self._h, i, buf, 128, ctypes.byref(ms),
                                                ctypes.byref(cnt), ctypes.byref(b)), "profile_get")
            out[buf.value.decode()] = (ms.value, cnt.value, b.value)
        return out

    def last_elapsed_ms(self) -> float:
        v = ctypes.c_float()
        _check(lib().accfft_amd_last_elapsed_ms(self._h, ctypes.byref(v)), "last_elapsed_ms")
        return v.value


# ------------------------------------------------------------------ reference-named API
def accfft_local_size_dft_r2c_gpu(n):
    """FP64 sizes (src/accfft_gpu.cpp:62-65); returns (alloc_max, isize, istart, osize, ostart)."""
    return local_size_dft_r2c(n, F64)


def accfft_local_size_dft_r2c_gpuf(n):
    return local_size_dft_r2c(n, F32)


def accfft_local_size_dft_c2c_gpu(n):
    return local_size_dft_c2c(n, F64)


def accfft_local_size_dft_c2c_gpuf(n):
    return local_size_dft_c2c(n, F32)


def accfft_plan_dft_3d_r2c_gpu(n, data, data_out, flags=ACCFFT_MEASURE) -> Plan:
    return Plan("r2c", n, data, data_out, F64, flags)


def accfft_plan_dft_3d_r2c_gpuf(n, data, data_out, flags=ACCFFT_MEASURE) -> Plan:
    return Plan("r2c", n, data, data_out, F32, flags)


def accfft_plan_dft_3d_c2c_gpu(n, data, data_out, flags=ACCFFT_MEASURE) -> Plan:
    return Plan("c2c", n, data, data_out, F64, flags)


def accfft_plan_dft_3d_c2c_gpuf(n, data, data_out, flags=ACCFFT_MEASURE) -> Plan:
    return Plan("c2c", n, data, data_out, F32, flags)


def accfft_execute_r2c_gpu(plan: Plan, data=None, data_out=None, timer=None, xyz=7):
    plan.execute_r2c(data, data_out, timer, xyz)


accfft_execute_r2c_gpuf = accfft_execute_r2c_gpu


def accfft_execute_c2r_gpu(plan: Plan, data=None, data_out=None, timer=None, xyz=7):
    plan.execute_c2r(data, data_out, timer, xyz)


accfft_execute_c2r_gpuf = accfft_execute_c2r_gpu


def accfft_execute_c2c_gpu(plan: Plan, direction, data=None, data_out=None, timer=None, xyz=7):
    plan.execute_c2c(direction, data, data_out, timer, xyz)


accfft_execute_c2c_gpuf = accfft_execute_c2c_gpu


def accfft_destroy_plan_gpu(plan: Plan):
    plan.destroy()


accfft_destroy_plan = accfft_destroy_plan_gpu


def accfft_cleanup_gpu():
    """src/accfft_gpu.cpp:45-47 (nothing to release)."""


accfft_cleanup_gpuf = accfft_cleanup_gpu


def accfft_grad_gpu(A_x, A_y, A_z, A, plan: Plan, pXYZ: int = 7, timer=None):
    plan.grad(A, A_x, A_y, A_z, pXYZ, timer)


accfft_grad_gpuf = accfft_grad_gpu


def accfft_laplace_gpu(LA, A, plan: Plan, timer=None):
    plan.laplace(A, LA, timer)


accfft_laplace_gpuf = accfft_laplace_gpu


def accfft_biharmonic_gpu(BA, A, plan: Plan, timer=None):
    plan.biharmonic(A, BA, timer)


accfft_biharmonic_gpuf = accfft_biharmonic_gpu


def accfft_divergence_gpu(divA, A_x, A_y, A_z, plan: Plan, timer=None):
    plan.divergence(A_x, A_y, A_z, divA, timer)


accfft_divergence_gpuf = accfft_divergence_gpu


def accfft_inv_laplace_gpu(invLA, A, plan: Plan, timer=None):
    plan.inv_laplace(A, invLA, timer)


accfft_inv_laplace_gpuf = accfft_inv_laplace_gpu


# ------------------------------------------------------------------ element-wise / transposes
def mult_wave_number(kind: int, wA, A, n, osize, ostart, scale_xyz: int = 7):
    """The extern "C" multipliers of src/operators_cuda.cu:421-486."""
    nb = osize[0] * osize[1] * osize[2] * 2 * (4 if _fp_of(A.dtype) == F32 else 8)
    if wA.dtype != A.dtype:
        raise AccfftError(f"mult_wave_number: dtypes {wA.dtype} and {A.dtype} differ")
    _check(lib().accfft_amd_mult_wave_number(kind, _fp_of(A.dtype), _gpu(wA, "wA", nb),
                                             _gpu(A, "A", nb), _i3(*n),
                                             _i3(*osize), _i3(*ostart), scale_xyz, _stream()),
           "mult_wave_number")


def poisson_fourier_filter_gpu(data_hat, N, isize, istart, method: int = 0):
    """poisson_fourier_filter_gpu(data_hat, N, isize, istart, methodNb),
    steps/step4/kernels.cu:113-146, in place on a half spectrum block."""
    nb = isize[0] * isize[1] * isize[2] * 2 * (4 if _fp_of(data_hat.dtype) == F32 else 8)
    _check(lib().accfft_amd_poisson_filter(_fp_of(data_hat.dtype), _gpu(data_hat, "data_hat", nb),
                                           _i3(*N),
                                           _i3(*isize), _i3(*istart), int(method), _stream()),
           "poisson_filter")


def axpy(alpha: float, x, y):
    """daxpy_gpu_c(f), src/operators_cuda.cu:470-474: y += alpha * x."""
    if x.dtype != y.dtype or y.numel() < x.numel():
        raise AccfftError(f"axpy: y ({y.dtype}, {y.numel()}) cannot take x ({x.dtype}, {x.numel()})")
    _check(lib().accfft_amd_axpy(_fp_of(x.dtype), x.numel(), alpha, _gpu(x, "x"), _gpu(y, "y"),
                                 _stream()),
           "axpy")


def local_transpose(r, c, t, inp, out):
    """local_transpose_cuda(r,c,t,in,out), src/transpose_cuda.cu:242-257."""
    nb = r * c * t * inp.element_size()
    _check(lib().accfft_amd_local_transpose(r, c, t, _gpu(inp, "in", nb), _gpu(out, "out", nb),
                                            inp.element_size(),
                                            _stream()), "local_transpose")


def local_transpose_ragged(r, c, t, t2, inp, out):
    """local_transpose_cuda(r,c,t,t2,in,out), src/transpose_cuda.cu:207-224."""
    _check(lib().accfft_amd_local_transpose_ragged(r, c, t, t2, _gpu(inp, "in"), _gpu(out, "out"),
                                                   inp.element_size(), _stream()),
           "local_transpose_ragged")


def local_transpose_col(r, c, t, t2, inp, out):
    """local_transpose_col_cuda(r,c,t,t2,in,out), src/transpose_cuda.cu:78-95."""
    _check(lib().accfft_amd_local_transpose_col(r, c, t, t2, _gpu(inp, "in"), _gpu(out, "out"),
                                                inp.element_size(), _stream()),
           "local_transpose_col")


def local_transpose_inplace(r, c, t, A):
    """local_transpose_cuda(r,c,t,A), src/transpose_cuda.cu:227-239."""
    _check(lib().accfft_amd_local_transpose_inplace(r, c, t, _gpu(A, "A", r * c * t * A.element_size()),
                                                    A.element_size(),
                                                    _stream()), "local_transpose_inplace")


def _i64arr(v):
    v = [int(x) for x in v]
    return (ctypes.c_int64 * max(len(v), 1))(*v)


def memcpy_v1_h1(nprocs_1, howmany, local_n0, n_tuples, local_n1_proc, send_recv, data, idist,
                 N1, local_1_start_proc):
    """Pencil pack, src/transpose_cuda.cu:97-143 (see accfft_amd_pencil_pack)."""
    _check(lib().accfft_amd_pencil_pack(nprocs_1, howmany, local_n0, n_tuples,
                                        _i64arr(local_n1_proc), _gpu(send_recv, "send_recv"),
                                        _gpu(data, "data"), idist,
                                        N1, _i64arr(local_1_start_proc), data.element_size(),
                                        _stream()), "pencil_pack")


def memcpy_v1_h2(nprocs_0, howmany, local_0_start_proc, local_n0_proc, data, odist, local_n1,
                 n_tuples, send_recv):
    """Pencil unpack, src/transpose_cuda.cu:145-197 (see accfft_amd_pencil_unpack)."""
    _check(lib().accfft_amd_pencil_unpack(nprocs_0, howmany, _i64arr(local_0_start_proc),
                                          _i64arr(local_n0_proc), _gpu(data, "data"), odist,
                                          local_n1, n_tuples, _gpu(send_recv, "send_recv"),
                                          data.element_size(),
                                          _stream()), "pencil_unpack")
