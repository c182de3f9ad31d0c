"""Seeded synthetic input fields of the benchmark workloads (numpy, FP64).

Every field lives on the uniform grid x_d = 2 pi i_d / N_d of [0, 2 pi)^3 in
the reference's layout (row-major ``[N0][N1][N2]``, x slowest).  Seeds
default to ``SEED0 + workload index`` (workload 0: the 64^3 plain method,
1: the 256^3 / 512^3 drop-in runs, 2: the split-precision runs, 3: the
differentiable loss, 4: the slab share and 1024^3).

- ``sines_noise``: sin x + sin y + sin z plus N(0, sigma) noise (the plain
  method's field; the reference's step programs use analytic fields of this
  kind, steps/step1/step1_gpu.cpp:135-145);
- ``normal``: i.i.d. N(0, 1);
- ``band_limited``: a random-phase field whose spectrum has amplitude |k|^p
  on the shell 1 <= |k| <= kmax (signed integer wavenumbers, k > N/2 taken
  as k - N), built as a half spectrum, inverted with a real 3-D FFT (which
  makes it Hermitian) and scaled to unit RMS;
- ``dynamic_range``: N(0, 1) * 10^U(-4, 4) per point.

These are plain library routines (numpy's generator, scipy.fft on the host
cores); nothing here runs on the GPU or touches the transforms under test.
"""
from __future__ import annotations

import os
from typing import Sequence

import numpy as np

SEED0 = 20261202


def _n3(n: Sequence[int]):
    n = tuple(int(v) for v in n)
    if len(n) != 3 or min(n) < 1:
        raise ValueError(f"bad grid {n}")
    return n


def _workers() -> int:
    try:
        return max(1, min(16, len(os.sched_getaffinity(0))))
    except AttributeError:  # pragma: no cover
        return max(1, min(16, os.cpu_count() or 1))


def normal(n, seed: int) -> np.ndarray:
    return np.random.default_rng(seed).standard_normal(_n3(n))


def dynamic_range(n, seed: int) -> np.ndarray:
    g = np.random.default_rng(seed)
    n = _n3(n)
    return g.standard_normal(n) * 10.0 ** g.uniform(-4.0, 4.0, n)


def sines_noise(n, seed: int, sigma: float = 1e-3) -> np.ndarray:
    n = _n3(n)
    t = [np.sin(2 * np.pi * np.arange(m) / m) for m in n]
    f = t[0][:, None, None] + t[1][None, :, None] + t[2][None, None, :]
    return f + sigma * np.random.default_rng(seed).standard_normal(n)


def _signed(m: int) -> np.ndarray:
    k = np.arange(m, dtype=np.float64)
    k[k > m // 2] -= m
    return k


def band_limited(n, seed: int, slope: float = -5.0 / 3.0, kmin: float = 1.0,
                 kmax: float = None) -> np.ndarray:
    """Random-phase field with |X(k)| = |k|^slope for kmin <= |k| <= kmax
    (default kmax = min(N)/4), unit RMS."""
    import scipy.fft

    n = _n3(n)
    if kmax is None:
        kmax = min(n) / 4.0
    nh = n[2] // 2 + 1
    kz = np.arange(nh, dtype=np.float64)
    k2 = (_signed(n[0])[:, None, None] ** 2 + _signed(n[1])[None, :, None] ** 2
          + (kz ** 2)[None, None, :])
    kabs = np.sqrt(k2)
    del k2
    inside = (kabs >= kmin) & (kabs <= kmax)
    amp = np.zeros_like(kabs)
    amp[inside] = kabs[inside] ** slope
    del kabs, inside
    phase = np.random.default_rng(seed).uniform(0.0, 2.0 * np.pi, amp.shape)
    spec = amp * np.exp(1j * phase)
    del amp, phase
    f = scipy.fft.irfftn(spec, s=n, workers=_workers())
    del spec
    rms = float(np.sqrt(np.mean(f * f)))
    if rms == 0.0:
        raise ValueError("empty band")
    f /= rms
    return f


def complex_of(make, n, seed: int, **kw) -> np.ndarray:
    """A complex field whose real and imaginary parts are two draws of
    ``make`` (seeds ``seed`` and ``seed + 1000``)."""
    return make(n, seed, **kw) + 1j * make(n, seed + 1000, **kw)


def checksum(a: np.ndarray):
    """(sum, sum of squares) in FP64, recorded beside a generated input."""
    a = np.asarray(a)
    if np.iscomplexobj(a):
        return (complex(a.sum()), float(np.vdot(a, a).real))
    return (float(a.sum()), float(np.dot(a.ravel(), a.ravel())))
