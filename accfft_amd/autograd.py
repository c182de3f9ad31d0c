"""torch.autograd bindings of the r2c plan's transforms and spectral operators.

Every map here is real-linear, so each backward is one call of the matching
adjoint entry point of the C ABI (``accfft_amd_adjoint_*``) or, for the
self-adjoint operators, of the forward entry point itself:

    r2c^T(G) = c2r(w G)         w = 1/2 on z bins other than DC / Nyquist
    c2r^T(y) = r2c(y) / w
    grad^T   = -div             div^T = -grad
    laplace, biharmonic, inv_laplace, inv_biharmonic: self-adjoint

(inner products: Re sum conj(a) b on half spectra, the dot product on real
fields).  The forward maps are those of include/accfft_operators_gpu.h:44-67
and src/operators_gpu.txx; grad and divergence use the per-axis spectral
derivative (src/operators_gpu.txx:32-118, 336-446 with each axis
differentiated by its own 1-D transform pair).

No forward saves a tensor for its backward (a linear map's derivative does
not depend on the point), so a chain of these functions keeps no activations
alive and recomputation-based checkpointing
(``torch.utils.checkpoint``) costs nothing extra.  The only intermediate
spectra are the plans' workspaces, and they are dropped when memory asks
for it (``set_memory_policy``): after a call when the device's free memory
is below a low-water mark (or after every call), and, when an allocation
fails, every plan's workspace and torch's cached blocks are released and
the call is repeated once.  ``Spectral.release`` drops them by hand.
"""
from __future__ import annotations

import weakref
from typing import Callable, Dict, Optional, Sequence, Tuple

import torch

from . import ACCFFT_MEASURE, ERR_ALLOC, F32, F64, PRESERVE_INPUT, AccfftError, Plan

__all__ = ["Spectral", "rfft3", "irfft3", "grad", "divergence", "laplace", "inv_laplace",
           "biharmonic", "inv_biharmonic", "set_memory_policy", "release_all"]

# When the plans' workspaces (the intermediate spectra) are dropped:
#   low_water: free device memory (bytes) below which a plan releases its
#              workspace after a call; None disables the check
#   keep:      False releases after every call (smallest footprint; each
#              call allocates again)
_POLICY = {"low_water": 1 << 30, "keep": True}
_LIVE: "weakref.WeakSet[Spectral]" = weakref.WeakSet()


def set_memory_policy(low_water: Optional[int] = 1 << 30, keep: bool = True) -> None:
    """Sets when intermediate spectra are dropped (see ``_POLICY``)."""
    if low_water is not None and low_water < 0:
        raise AccfftError("low_water must be >= 0 or None")
    _POLICY["low_water"] = low_water
    _POLICY["keep"] = bool(keep)


def release_all() -> None:
    """Releases the workspace of every live Spectral."""
    for s in list(_LIVE):
        s.release()


def _retry_on_oom(fn: Callable, *args):
    """Runs fn; when a device allocation fails (the library's or torch's),
    drops every plan's workspace and torch's cached blocks and runs it once
    more."""
    try:
        return fn(*args)
    except AccfftError as e:
        if e.status != ERR_ALLOC:
            raise
    except torch.cuda.OutOfMemoryError:
        pass
    release_all()
    torch.cuda.empty_cache()
    return fn(*args)


class Spectral:
    """Plans for one grid shape and precision, on the current device.

    The plan is out of place; it is always called with explicit buffers, so
    the placeholders it was created with are never read or written."""

    def __init__(self, n: Sequence[int], dtype=torch.float32, device=None, flags: int = 0):
        if dtype not in (torch.float32, torch.float64):
            raise AccfftError(f"unsupported dtype {dtype}")
        self.n = tuple(int(v) for v in n)
        self.dtype = dtype
        self.cdtype = torch.complex64 if dtype == torch.float32 else torch.complex128
        self.device = torch.device(device if device is not None else "cuda")
        self._ph = (torch.empty(1, dtype=dtype, device=self.device),
                    torch.empty(1, dtype=dtype, device=self.device))
        prec = F32 if dtype == torch.float32 else F64
        self.plan = _retry_on_oom(lambda: Plan("r2c", self.n, self._ph[0], self._ph[1], prec,
                                               ACCFFT_MEASURE | PRESERVE_INPUT | flags))
        self.releases = 0  # workspace releases made by the memory policy
        _LIVE.add(self)

    def _call(self, fn: Callable, *args):
        """One plan call under the memory policy."""
        _retry_on_oom(fn, *args)
        if not _POLICY["keep"]:
            self._drop()
            return
        lw = _POLICY["low_water"]
        if lw is not None and self.plan.workspace_bytes() > 0:
            free, _ = torch.cuda.mem_get_info(self.device)
            if free < lw:
                self._drop()

    def _drop(self):
        if self.plan.workspace_bytes() > 0:
            self.release()
            self.releases += 1

    @property
    def spectrum_shape(self) -> Tuple[int, int, int]:
        return (self.n[0], self.n[1], self.n[2] // 2 + 1)

    def release(self):
        self.plan.release_workspace()

    def _real(self, x: torch.Tensor) -> torch.Tensor:
        if tuple(x.shape) != self.n or x.dtype != self.dtype or x.device != self.device:
            raise AccfftError(f"expected {self.dtype} {self.n} on {self.device}, got "
                              f"{x.dtype} {tuple(x.shape)} on {x.device}")
        return x.contiguous()

    def _spec(self, G: torch.Tensor) -> torch.Tensor:
        if (tuple(G.shape) != self.spectrum_shape or G.dtype != self.cdtype
                or G.device != self.device):
            raise AccfftError(f"expected {self.cdtype} {self.spectrum_shape} on {self.device}, "
                              f"got {G.dtype} {tuple(G.shape)} on {G.device}")
        return G.contiguous()

    def new_real(self):
        return _retry_on_oom(lambda: torch.empty(self.n, dtype=self.dtype, device=self.device))

    def new_spec(self):
        return _retry_on_oom(lambda: torch.empty(self.spectrum_shape, dtype=self.cdtype,
                                                 device=self.device))

    # raw (non-differentiable) calls
    def r2c(self, x):
        G = self.new_spec()
        self._call(self.plan.execute_r2c, self._real(x), G)
        return G

    def c2r(self, G):
        y = self.new_real()
        self._call(self.plan.execute_c2r, self._spec(G), y)  # PRESERVE_INPUT: G is kept
        return y

    def r2c_adjoint(self, G):
        x = self.new_real()
        self._call(self.plan.adjoint_r2c, self._spec(G), x)
        return x

    def c2r_adjoint(self, y):
        G = self.new_spec()
        self._call(self.plan.adjoint_c2r, self._real(y), G)
        return G

    def grad(self, u):
        outs = [self.new_real() for _ in range(3)]
        self._call(self.plan.grad, self._real(u), *outs)
        return tuple(outs)

    def grad_adjoint(self, gx, gy, gz):
        out = self.new_real()
        self._call(self.plan.adjoint_grad,
                   *(None if g is None else self._real(g) for g in (gx, gy, gz)), out)
        return out

    def divergence(self, vx, vy, vz):
        out = self.new_real()
        self._call(self.plan.divergence, self._real(vx), self._real(vy), self._real(vz), out)
        return out

    def divergence_adjoint(self, v, need=(True, True, True)):
        outs = [self.new_real() if k else None for k in need]
        self._call(self.plan.adjoint_divergence, self._real(v), *outs)
        return tuple(outs)

    def op(self, name: str, u):
        out = self.new_real()
        self._call(getattr(self.plan, name), self._real(u), out)
        return out


_CTX: Dict[tuple, Spectral] = {}


def context(n, dtype, device) -> Spectral:
    """The cached Spectral for (shape, dtype, device, current stream).

    One plan per stream: a plan's calls share its workspaces, and the plan
    orders calls that arrive on different streams one after the other, so
    work on two streams gets plans of its own to overlap."""
    dev = torch.device(device)
    stream = torch.cuda.current_stream(dev).cuda_stream if dev.type == "cuda" else 0
    key = (tuple(int(v) for v in n), dtype, dev, stream)
    s = _CTX.get(key)
    if s is None:
        s = Spectral(key[0], dtype, dev)
        _CTX[key] = s
    return s


class _R2C(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, S):
        ctx.S = S
        return S.r2c(x)

    @staticmethod
    def backward(ctx, G):
        # torch's gradient convention for a real input of a complex output:
        # dL/dx = Re sum conj(dL/dG) dG/dx, i.e. the real-linear adjoint
        return ctx.S.r2c_adjoint(G), None


class _C2R(torch.autograd.Function):
    @staticmethod
    def forward(ctx, G, S):
        ctx.S = S
        return S.c2r(G)

    @staticmethod
    def backward(ctx, y):
        return ctx.S.c2r_adjoint(y), None


class _Grad(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u, S):
        ctx.S = S
        return S.grad(u)

    @staticmethod
    def backward(ctx, gx, gy, gz):
        return ctx.S.grad_adjoint(gx, gy, gz), None


class _Div(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vx, vy, vz, S):
        ctx.S = S
        return S.divergence(vx, vy, vz)

    @staticmethod
    def backward(ctx, v):
        need = ctx.needs_input_grad[:3]
        return ctx.S.divergence_adjoint(v, need) + (None,)


class _SelfAdjoint(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u, S, name):
        ctx.S, ctx.name = S, name
        return S.op(name, u)

    @staticmethod
    def backward(ctx, g):
        return ctx.S.op(ctx.name, g), None, None


def _ctx_of(x: torch.Tensor, S: Optional[Spectral]) -> Spectral:
    return S if S is not None else context(x.shape, x.dtype, x.device)


def rfft3(x: torch.Tensor, S: Optional[Spectral] = None) -> torch.Tensor:
    """Unnormalised forward r2c transform (sign -1), half spectrum along z."""
    return _R2C.apply(x, _ctx_of(x, S))


def irfft3(G: torch.Tensor, S: Spectral) -> torch.Tensor:
    """Unnormalised backward c2r transform (sign +1); ``G`` is not modified.
    The Spectral is required: a half spectrum does not fix the z length."""
    return _C2R.apply(G, S)


def grad(u: torch.Tensor, S: Optional[Spectral] = None):
    return _Grad.apply(u, _ctx_of(u, S))


def divergence(vx, vy, vz, S: Optional[Spectral] = None):
    return _Div.apply(vx, vy, vz, _ctx_of(vx, S))


def laplace(u, S: Optional[Spectral] = None):
    return _SelfAdjoint.apply(u, _ctx_of(u, S), "laplace")


def inv_laplace(u, S: Optional[Spectral] = None):
    return _SelfAdjoint.apply(u, _ctx_of(u, S), "inv_laplace")


def biharmonic(u, S: Optional[Spectral] = None):
    return _SelfAdjoint.apply(u, _ctx_of(u, S), "biharmonic")


def inv_biharmonic(u, S: Optional[Spectral] = None):
    return _SelfAdjoint.apply(u, _ctx_of(u, S), "inv_biharmonic")
