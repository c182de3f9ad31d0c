#!/usr/bin/env python
"""Headline benchmark: 512^3 FP32 r2c + c2r round trip on one MI355X.

One "step" is one forward r2c (z R2C, y C2C, x C2C) plus one inverse c2r
(x C2C, y C2C, z C2R) of a 512x512x512 FP32 field already resident in HBM,
through the C ABI of libaccfft_amd.so (the path of
accfft_execute_r2c_gpuf / accfft_execute_c2r_gpuf, src/accfft_gpuf.cpp:443-675).

Prints ONE JSON line.  ``value`` is milliseconds per round trip (lower is
better), the time form of the metric; ``roofline`` gives the dominant
kernel's achieved algorithmic GB/s against the MI355X HBM peak, measured
with HIP events around each launch on the plan's stream during a timed
pass, and its HBM bytes per launch from the committed PMC summary when that
summary was made from the current kernel sources; ``cpu_baseline`` gives
the host's CPU timings (median of 5 after 2 warm-ups): the FP64 CPU oracle
(oracle/oracle.c, OpenMP) on the same round trip, and torch.fft on the CPU
in FP32 / FP64 at 64^3 .. 512^3 (FFTW, the reference's CPU library, is
probed and absent on this image).  The split-precision leg runs the
workload's three seeded inputs and reports each form's rel-L2 and max-abs
error against the FP64 oracle.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def alg_bytes_roundtrip(n, s):
    """Pass model of the reference algorithm: r2c = R + 5C, c2r = 5C + R."""
    N0, N1, N2 = n
    R = N0 * N1 * N2 * s
    C = N0 * N1 * (N2 // 2 + 1) * 2 * s
    return 2 * R + 10 * C, 2 * (R + C)


ROOT = os.path.dirname(os.path.abspath(__file__))


def kernel_sources_sha256():
    """Hash of the HIP sources (accfft_amd/csrc, sorted by name): a PMC
    summary stamped with another hash was made from other kernels."""
    import hashlib
    h = hashlib.sha256()
    d = os.path.join(ROOT, "accfft_amd", "csrc")
    for name in sorted(os.listdir(d)):
        if name.endswith((".hip", ".hpp", ".inc", ".h")):
            h.update(name.encode())
            with open(os.path.join(d, name), "rb") as f:
                h.update(f.read())
    return h.hexdigest()


def profiled_traffic(kernel, n):
    """HBM bytes per launch of ``kernel`` from the committed rocprofv3 PMC
    summary (FETCH_SIZE + WRITE_SIZE, scaled by kernels of known size) and
    where they come from.  The bytes are None when no summary covers the
    kernel or the summary's source stamp differs from the current kernel
    sources (the counters would describe other code)."""
    path = os.path.join(ROOT, "profiles", f"fft_passes_{n[0]}cubed_fp32_hbm_traffic.md")
    src = {"file": os.path.relpath(path, ROOT), "sources_sha256": None, "current": False}
    if n[0] != n[1] or n[1] != n[2] or not os.path.exists(path):
        return None, src
    val = None
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) == 2 and parts[0] == "sources_sha256:":
                src["sources_sha256"] = parts[1]
            if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch:" and parts[1] == kernel:
                val = int(parts[2])
    src["current"] = src["sources_sha256"] == kernel_sources_sha256()
    return (val if src["current"] else None), src


def run_gpu(n, fp, steps, warmup, profile=True):
    import torch
    import accfft_amd as af

    dev = torch.device("cuda:0")
    rdt = torch.float32 if fp == af.F32 else torch.float64
    cdt = torch.complex64 if fp == af.F32 else torch.complex128
    g = torch.Generator(device=dev)
    g.manual_seed(20261203)
    x = torch.randn(n, generator=g, device=dev, dtype=rdt)
    spec = torch.empty((n[0], n[1], n[2] // 2 + 1), device=dev, dtype=cdt)
    back = torch.empty(n, device=dev, dtype=rdt)
    plan = af.Plan("r2c", n, x, spec, fp)  # ACCFFT_MEASURE: tiles tuned at plan time
    tuning = list(plan.tuning())

    def step():
        plan.execute_r2c(x, spec)
        plan.execute_c2r(spec, back)  # overwrites spec, as the reference

    stream = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(steps):
        step()
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    # correctness of the timed configuration: the last round trip = N^3 * x
    # (checked after the timed region: a host round trip before it leaves
    # the GPU idle and slows the first timed steps)
    err = (torch.linalg.vector_norm(back / float(n[0] * n[1] * n[2]) - x) /
           torch.linalg.vector_norm(x)).item()

    # practical HBM ceiling on this box: a device-to-device copy of the
    # same footprint (read + write), for context next to the FFT passes
    src = spec.view(torch.float32) if fp == af.F32 else spec.view(torch.float64)
    dst = torch.empty_like(src)
    for _ in range(2):
        dst.copy_(src)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(5):
        dst.copy_(src)
    e1.record(stream)
    torch.cuda.synchronize()
    copy_gbps = 2 * src.numel() * src.element_size() / (e0.elapsed_time(e1) / 5 * 1e-3) / 1e9
    del dst

    # the same round trip with the z pass and the y transform's first level
    # as separate passes (tuning choice -3 on the y axis), for comparison
    unfused_ms = None
    if tuning[1] != -3:
        try:
            plan.set_tuning(1, -3)
            fused_now = plan.tuning()[1] == -3
        except Exception:
            fused_now = False
        if fused_now:
            for _ in range(warmup):
                step()
            torch.cuda.synchronize()
            e0.record(stream)
            for _ in range(steps):
                step()
            e1.record(stream)
            torch.cuda.synchronize()
            unfused_ms = e0.elapsed_time(e1) / steps
            plan.set_tuning(1, -1)

    prof = {}
    if profile:
        plan.set_profiling(True)
        plan.profile_reset()
        for _ in range(steps):
            step()
        torch.cuda.synchronize()
        prof = plan.profile()
        plan.set_profiling(False)
    plan.destroy()
    return ms, err, prof, copy_gbps, tuning, unfused_ms


SPLIT_FLAGS = {"bf16x2": 0x1000, "bf16x3": 0x2000, "fp16x2": 0x4000}
SPLIT_DISTS = ("normal", "band_limited", "dynamic_range")


def split_inputs(n, dist):
    """The split-precision workload's seeded inputs (accfft_amd/fields.py):
    a real field for the r2c and a complex one for the c2c, FP32."""
    import numpy as np
    from accfft_amd import fields
    make = {"normal": fields.normal, "dynamic_range": fields.dynamic_range,
            "band_limited": lambda m, s: fields.band_limited(m, s, slope=-5.0 / 3.0)}[dist]
    seed = fields.SEED0 + 2
    x = make(n, seed).astype(np.float32)
    z = fields.complex_of(make, n, seed).astype(np.complex64)
    return x, z


def run_split(n, steps, dists=SPLIT_DISTS):
    """512^3 FP32 forward c2c and r2c: VALU butterflies vs split-precision
    matrix-core radix-16 stages, timed on the N(0,1) input.  Every form's
    output on each of the workload's three inputs is kept on the GPU so that
    the CPU leg can check it against the FP64 oracle (rel-L2 and max-abs)."""
    import torch
    import accfft_amd as af

    dev = torch.device("cuda:0")
    out, outputs = {}, {}

    def timeit(fn):
        for _ in range(2):
            fn()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps

    forms = [("fp32_valu", 0)] + list(SPLIT_FLAGS.items())
    h = torch.empty((n[0], n[1], n[2] // 2 + 1), dtype=torch.complex64, device=dev)
    for dist in dists:
        x, z = split_inputs(n, dist)
        xd = torch.from_numpy(x).to(dev)
        zd = torch.from_numpy(z).to(dev)
        del x, z
        zo = torch.empty_like(zd)
        for name, flag in forms:
            pc = af.Plan("c2c", n, zd, zo, af.F32, af.ACCFFT_MEASURE | flag)
            pr = af.Plan("r2c", n, xd, h, af.F32, af.ACCFFT_MEASURE | flag)
            if dist == "normal":
                tc = timeit(lambda: pc.execute_c2c(-1, zd, zo))
                tr = timeit(lambda: pr.execute_r2c(xd, h))
                out[name] = {"c2c_ms": round(tc, 4), "r2c_ms": round(tr, 4)}
                log(f"  {name:10s} c2c {tc:.4f} ms  r2c {tr:.4f} ms")
            pc.execute_c2c(-1, zd, zo)
            pr.execute_r2c(xd, h)
            torch.cuda.synchronize()
            outputs[(dist, "c2c", name)] = zo.clone()
            outputs[(dist, "r2c", name)] = h.clone()
            pc.destroy()
            pr.destroy()
        del xd, zd, zo
    # 256^3: the arrays (67 MB real, 135 MB complex) stay in the 256 MB
    # memory-side cache between passes, so HBM is not the binding roof there
    m = (256, 256, 256)
    g = torch.Generator(device=dev)
    g.manual_seed(20261207)
    xs = torch.randn(m, generator=g, device=dev)
    zs = torch.complex(torch.randn(m, generator=g, device=dev), torch.randn(m, generator=g, device=dev))
    hs = torch.empty((m[0], m[1], m[2] // 2 + 1), dtype=torch.complex64, device=dev)
    bs = torch.empty_like(xs)
    zo = torch.empty_like(zs)
    for name, flag in forms:
        pr = af.Plan("r2c", m, xs, hs, af.F32, af.ACCFFT_MEASURE | flag)
        pc = af.Plan("c2c", m, zs, zo, af.F32, af.ACCFFT_MEASURE | flag)

        def rt():
            pr.execute_r2c(xs, hs)
            pr.execute_c2r(hs, bs)
        out[name]["r2c_c2r_256_ms"] = round(timeit(rt), 4)
        out[name]["c2c_256_ms"] = round(timeit(lambda: pc.execute_c2c(-1, zs, zo)), 4)
        pr.destroy()
        pc.destroy()
        log(f"  {name:10s} 256^3 round trip {out[name]['r2c_c2r_256_ms']:.4f} ms  "
            f"c2c {out[name]['c2c_256_ms']:.4f} ms")
    base = out["fp32_valu"]
    for name in SPLIT_FLAGS:
        o = out[name]
        o["speedup_c2c_vs_fp32"] = round(base["c2c_ms"] / o["c2c_ms"], 4)
        o["speedup_r2c_vs_fp32"] = round(base["r2c_ms"] / o["r2c_ms"], 4)
        o["speedup_256_round_trip_vs_fp32"] = round(base["r2c_c2r_256_ms"] / o["r2c_c2r_256_ms"], 4)
        o["speedup_256_c2c_vs_fp32"] = round(base["c2c_256_ms"] / o["c2c_256_ms"], 4)
    return out, outputs


def check_split(n, split, outputs):
    """rel-L2 and max-abs error of every precision form against the FP64
    oracle (oracle/oracle.c) on the workload's three inputs; the oracle is
    the checker here, its outputs are compared, not timed."""
    import numpy as np
    import torch
    import oracle

    oracle.build()
    dev = torch.device("cuda:0")
    for dist in sorted({k[0] for k in outputs}, key=SPLIT_DISTS.index):
        x, z = split_inputs(n, dist)
        for kind in ("r2c", "c2c"):
            ref = (oracle.r2c(x.astype(np.float64)) if kind == "r2c"
                   else oracle.c2c(z.astype(np.complex128), -1))
            rd = torch.from_numpy(ref).to(dev)
            del ref
            nr = torch.linalg.vector_norm(rd).item()
            mr = rd.abs().max().item()
            for name in split:
                o = outputs.pop((dist, kind, name))
                d = o.to(torch.complex128) - rd
                split[name][f"rel_l2_{kind}_{dist}"] = torch.linalg.vector_norm(d).item() / nr
                split[name][f"max_abs_{kind}_{dist}"] = d.abs().max().item()
                split[name][f"max_abs_rel_{kind}_{dist}"] = d.abs().max().item() / mr
                del o, d
            del rd
        del x, z
    for name, o in split.items():
        worst = max(v for k, v in o.items() if k.startswith("rel_l2_"))
        o["rel_l2_max"] = worst
        o["meets_rel_err_1e-5"] = worst <= 1e-5
        log(f"  {name:10s} worst rel-L2 vs the FP64 oracle {worst:.2e}")


def _timeit(fn, steps, warmup=2):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def run_wide(steps):
    """The larger shapes: one GPU's 256 x 2048 x 2048 FP32 share of a 2048^3
    r2c on 8 ranks (exchange = identity), fused (pack in the y store) and in
    the reference's structure (z+y, separate pack kernel, x), and a 1024^3
    FP32 round trip, plain and split-BF16."""
    import torch
    import accfft_amd as af

    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(20261205)
    out = {}
    # ---- slab share
    N, P = (2048, 2048, 2048), 8
    n0, n1, Nh = N[0] // P, N[1] // P, N[2] // 2 + 1
    R = n0 * N[1] * N[2] * 4
    C = n0 * N[1] * Nh * 8
    u = torch.randn((n0, N[1], N[2]), generator=g, device=dev)
    h = torch.empty((N[0], n1, Nh), dtype=torch.complex64, device=dev)
    back = torch.empty_like(u)
    ps = af.Plan("r2c_slab", N, u, h, af.F32, af.ACCFFT_MEASURE, nprocs=P)
    t_fwd = _timeit(lambda: ps.execute_r2c(u, h), steps)

    def rt():
        ps.execute_r2c(u, h)
        ps.execute_c2r(h, back)
    t_rt = _timeit(rt, steps)
    err = (torch.linalg.vector_norm(back / float(N[0] * N[1] * N[2]) - u) /
           torch.linalg.vector_norm(u)).item()
    ps.destroy()
    # the reference's structure: z+y r2c on the slab, pack kernel, x c2c
    zy = torch.empty((n0, N[1], Nh), dtype=torch.complex64, device=dev)
    pl = af.Plan("r2c", (n0, N[1], N[2]), u, zy, af.F32)
    px = af.Plan("c2c", (N[0], n1, Nh), h, h, af.F32)

    def ref_fwd():
        pl.execute_r2c(u, zy, xyz=6)
        af.local_transpose_col(n0, P, n1 * Nh * 2, n1 * Nh * 2, zy.view(torch.float32), h.view(torch.float32))
        px.execute_c2c(-1, h, h, xyz=1)
    t_ref = _timeit(ref_fwd, steps)
    pl.destroy()
    px.destroy()
    del zy, back
    out["slab_256x2048x2048_p8"] = {
        "fwd_ms": round(t_fwd, 4), "roundtrip_ms": round(t_rt, 4), "roundtrip_rel_l2": err,
        "fwd_pass_model_bytes_fused": R + 5 * C,
        "fwd_frac_fused_model": round((R + 5 * C) / (t_fwd * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "fwd_frac_reference_model": round((R + 9 * C) / (t_fwd * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "reference_structure_fwd_ms": round(t_ref, 4),
    }
    log(f"  slab 256x2048x2048 P=8: fwd {t_fwd:.3f} ms (ref structure {t_ref:.3f} ms), "
        f"round trip {t_rt:.3f} ms, rel err {err:.2e}")
    del u, h
    torch.cuda.empty_cache()
    out.update(run_pencil_share(steps, g))
    torch.cuda.empty_cache()
    # ---- 1024^3 round trip, plain and split-BF16
    n = (1024, 1024, 1024)
    x = torch.randn(n, generator=g, device=dev)
    spec = torch.empty((n[0], n[1], n[2] // 2 + 1), dtype=torch.complex64, device=dev)
    back = torch.empty_like(x)
    alg, _ = alg_bytes_roundtrip(n, 4)
    for name, flag in (("fp32_valu", 0), ("bf16x2", SPLIT_FLAGS["bf16x2"])):
        p = af.Plan("r2c", n, x, spec, af.F32, af.ACCFFT_MEASURE | flag)

        def step():
            p.execute_r2c(x, spec)
            p.execute_c2r(spec, back)
        t = _timeit(step, max(3, steps // 2))
        e = (torch.linalg.vector_norm(back / float(n[0] * n[1] * n[2]) - x) /
             torch.linalg.vector_norm(x)).item()
        p.destroy()
        out[f"r2c_c2r_1024_{name}"] = {"ms": round(t, 4), "roundtrip_rel_l2": e,
                                       "pass_model_frac": round(alg / (t * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        log(f"  1024^3 {name}: round trip {t:.3f} ms, rel err {e:.2e}")
    return out


def run_pencil_share(steps, g):
    """One GPU's share of a 2048^3 FP32 r2c on a 4 x 2 pencil grid, rank
    (0, 0): [512][1024][2048] real -> [2048][512][513] complex (the half
    spectrum's 1025 z columns split 513 / 512, so the row transposes are
    ragged; the y rows split in equal power-of-two blocks, so the packs and
    unpacks ride on the z and y passes).  The exchange hook returns at once
    (the other ranks live on other GPUs; the received blocks hold stale
    data), so the time is the local part the reference reports in
    timer[4]: z r2c into the row exchange's blocks, y c2c from the received
    blocks into the column exchange's blocks, x c2c (and the inverse)."""
    import torch
    import accfft_amd as af

    dev = torch.device("cuda:0")
    N, grid = (2048, 2048, 2048), (4, 2)
    p = af.Plan("r2c_pencil", N, None, None, af.F32, af.ACCFFT_ESTIMATE, grid=grid,
                coords=(0, 0))
    n0, n1, _ = p.isize
    _, ny, nz = p.osize
    Nh = N[2] // 2 + 1
    u = torch.randn((n0, n1, N[2]), generator=g, device=dev)
    h = torch.empty(p.osize, dtype=torch.complex64, device=dev)
    back = torch.empty_like(u)
    p.set_exchange_v(lambda *a: 0)
    t_fwd = _timeit(lambda: p.execute_r2c(u, h), steps)
    t_inv = _timeit(lambda: p.execute_c2r(h, back), steps)
    p.destroy()
    R = n0 * n1 * N[2] * 4
    C1 = n0 * n1 * Nh * 8  # [n0][n1][Nh]
    C2 = n0 * N[1] * nz * 8  # [n0][N1][nz] = [N0][ny][nz]
    # the z, y and x passes with the packs and unpacks in their stores and
    # loads (equal power-of-two y blocks over both communicators)
    model = R + C1 + 4 * C2
    copies = R + 3 * C1 + 8 * C2  # with separate pack / unpack copies, as the reference
    res = {"pencil_2048_4x2_rank00_f32": {
        "fwd_ms": round(t_fwd, 4), "inv_ms": round(t_inv, 4),
        "fwd_pass_model_bytes": model,
        "fwd_frac_pass_model": round(model / (t_fwd * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "fwd_frac_separate_copies_model": round(copies / (t_fwd * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "exchange": "excluded (hook returns at once)"}}
    log(f"  pencil 2048^3 4x2 rank (0,0) [{n0}][{n1}][{N[2]}] -> [{N[0]}][{ny}][{nz}]: "
        f"fwd {t_fwd:.3f} ms, inv {t_inv:.3f} ms")
    return res


def run_ops(steps):
    """The spectral operators at 512^3 FP32 (accfft_grad_gpu with 3 outputs,
    laplace, inverse laplace, biharmonic, divergence, the step4 Poisson
    solve) and the differentiable loss 1/2 |div(grad(invlap u)) - g|^2 at
    256^3 FP32 with its adjoint backward.  Bytes are this library's passes
    (each operand read and written once per pass): grad 3 x 2R (one fused
    per-axis derivative pass per component), the spectral operators 2R + 8C
    (z r2c, y, x forward + symbol + inverse in one pass, y, z c2r),
    divergence 8R (2R + two accumulating passes of 3R)."""
    import torch
    import accfft_amd as af
    from accfft_amd import autograd as ag

    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(20261206)
    n = (512, 512, 512)
    R = n[0] * n[1] * n[2] * 4
    C = n[0] * n[1] * (n[2] // 2 + 1) * 8
    u = torch.randn(n, generator=g, device=dev)
    h = torch.empty((n[0], n[1], n[2] // 2 + 1), dtype=torch.complex64, device=dev)
    o = [torch.empty_like(u) for _ in range(3)]
    plan = af.Plan("r2c", n, u, h, af.F32)
    out = {}
    ops = [
        ("grad_3_components", lambda: plan.grad(u, o[0], o[1], o[2]), 6 * R),
        ("laplace", lambda: plan.laplace(u, o[0]), 2 * R + 8 * C),
        ("inv_laplace", lambda: plan.inv_laplace(u, o[0]), 2 * R + 8 * C),
        ("biharmonic", lambda: plan.biharmonic(u, o[0]), 2 * R + 8 * C),
        ("poisson_solve", lambda: plan.poisson_solve(u, o[0], 1), 2 * R + 8 * C),
    ]
    div = torch.empty_like(u)
    ops.append(("divergence", lambda: plan.divergence(o[0], o[1], o[2], div), 8 * R))
    for name, fn, b in ops:
        t = _timeit(fn, steps)
        out[name] = {"ms": round(t, 4), "alg_GBps": round(b / (t * 1e-3) / 1e9, 1),
                     "frac_of_hbm_peak": round(b / (t * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        log(f"  {name:18s} {t:.3f} ms  {b / (t * 1e-3) / 1e9:7.0f} GB/s")
    plan.destroy()
    del u, h, o
    torch.cuda.empty_cache()
    # 256^3 round trips (the arrays fit the 256 MB memory-side cache, so
    # HBM is not the binding roof at this size)
    m = (256, 256, 256)
    for prec, rdt, cdt in ((af.F32, torch.float32, torch.complex64),
                           (af.F64, torch.float64, torch.complex128)):
        xs = torch.randn(m, generator=g, device=dev, dtype=rdt)
        hs = torch.empty((m[0], m[1], m[2] // 2 + 1), dtype=cdt, device=dev)
        bs = torch.empty_like(xs)
        ps = af.Plan("r2c", m, xs, hs, prec)

        def rt256():
            ps.execute_r2c(xs, hs)
            ps.execute_c2r(hs, bs)
        t = _timeit(rt256, steps)
        name = "r2c_c2r_256_" + ("fp32" if prec == af.F32 else "fp64")
        Rm = m[0] * m[1] * m[2] * xs.element_size()
        Cm = m[0] * m[1] * (m[2] // 2 + 1) * hs.element_size()
        out[name] = {"ms": round(t, 4),
                     "pass_model_GBps": round((2 * Rm + 10 * Cm) / (t * 1e-3) / 1e9, 1)}
        log(f"  {name:18s} {t:.3f} ms")
        ps.destroy()
        del xs, hs, bs
    # differentiation: forward + backward of the loss at 256^3 FP32
    S = ag.context(m, torch.float32, dev)
    v = torch.randn(m, generator=g, device=dev, requires_grad=True)
    tgt = torch.randn(m, generator=g, device=dev)

    def fwd_bwd():
        v.grad = None
        r = ag.divergence(*ag.grad(ag.inv_laplace(v, S), S), S) - tgt
        loss = 0.5 * (r * r).sum()
        loss.backward()
    t = _timeit(fwd_bwd, steps)
    out["loss_fwd_bwd_256"] = {"ms": round(t, 4)}
    log(f"  loss fwd+bwd 256^3  {t:.3f} ms")
    return out


def _median_ms(fn, runs=5, warmup=2):
    """Median wall time of ``runs`` calls after ``warmup`` calls, in ms
    (the reference's practice: steps/step1/step1_gpu.cpp:57-58)."""
    import statistics
    for _ in range(warmup):
        fn()
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        ts.append((time.perf_counter() - t0) * 1e3)
    return statistics.median(ts)


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _probe_fftw():
    """The reference's CPU path is FFTW (src/accfft.cpp:179-338): probe for
    the libraries with dlopen."""
    import ctypes
    import ctypes.util
    found = {}
    for lib in ("fftw3", "fftw3f", "fftw3_threads", "fftw3f_threads"):
        path = ctypes.util.find_library(lib)
        ok = False
        for cand in ([path] if path else []) + [f"lib{lib}.so.3", f"lib{lib}.so"]:
            try:
                ctypes.CDLL(cand)
                ok = True
                break
            except OSError:
                continue
        found[lib] = ok
    return found


def cpu_baseline(n, full=False):
    """CPU timings on the host cores, median of 5 runs after 2 warm-ups.

    The reference's own CPU path (FFTW, 1 MPI rank + OpenMP) cannot run: the
    libraries are probed and, when absent, the comparator is a library CPU
    FFT, torch.fft.rfftn + irfftn (pocketfft) in FP32 and FP64, at 64^3,
    256^3, 512^3 (and 1024^3 FP32 with ``full``).  ``value`` is this
    build's FP64 oracle (oracle/oracle.c, OpenMP) on the headline's 512^3
    round trip, the "port" the contract names."""
    import numpy as np
    import torch
    import oracle

    try:
        threads = len(os.sched_getaffinity(0))
    except AttributeError:
        threads = os.cpu_count() or 1
    threads = max(1, min(threads, 16))
    fftw = _probe_fftw()
    torch.set_num_threads(threads)
    rng = np.random.default_rng(7)
    lib = {}
    sizes = [64, 256, 512] + ([1024] if full else [])
    for m in sizes:
        for dt, nm in ((torch.float32, "fp32"), (torch.float64, "fp64")):
            a = torch.from_numpy(rng.standard_normal((m, m, m))).to(dt)
            t = _median_ms(lambda: torch.fft.irfftn(torch.fft.rfftn(a), s=a.shape))
            lib[f"{m}^3_{nm}_ms"] = round(t, 3)
            log(f"  cpu torch.fft {m}^3 {nm} round trip {t:.2f} ms")
            del a
    oracle.build()
    oracle.set_threads(threads)
    a = rng.standard_normal(n)
    t_port = _median_ms(lambda: oracle.c2r(oracle.r2c(a), n))
    log(f"  cpu oracle {n} FP64 round trip {t_port:.1f} ms")
    return {"value": round(t_port, 3), "unit": "ms per round trip", "cores": threads,
            "kind": "port",
            "sample": (f"FP64 r2c+c2r round trip of {n[0]}x{n[1]}x{n[2]} N(0,1) with "
                       f"oracle/oracle.c (OpenMP, {threads} threads), median of 5 after 2 "
                       "warm-ups"),
            "cpu_model": _cpu_model(), "nproc": os.cpu_count(),
            "fftw_present": fftw,
            "library_torch_fft_round_trip": lib,
            "library_note": ("FFTW absent: torch.fft.rfftn+irfftn on the CPU (pocketfft, "
                             f"{threads} threads), median of 5 after 2 warm-ups"
                             if not any(fftw.values()) else "FFTW present; not driven")}


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true",
                    help="skip the CPU baseline and the oracle checks of the split leg")
    ap.add_argument("--cpu-full", action="store_true",
                    help="also time the CPU library FFT at 1024^3")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 side measurement")
    ap.add_argument("--no-split", action="store_true", help="skip the split-precision leg")
    ap.add_argument("--split-dists", default=",".join(SPLIT_DISTS),
                    help="inputs of the split-precision leg (comma-separated subset of "
                         + ",".join(SPLIT_DISTS) + ")")
    ap.add_argument("--no-wide", action="store_true",
                    help="skip the slab-share and 1024^3 legs")
    ap.add_argument("--no-ops", action="store_true",
                    help="skip the operator and differentiation leg")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the single-GPU "
              "hot path has no inter-GPU exchange in this build.", file=sys.stderr)
        sys.exit(2)

    import accfft_amd as af

    n = (args.n, args.n, args.n)
    ms, err, prof, copy_gbps, tuning, unfused_ms = run_gpu(n, af.F32, args.steps, args.warmup)
    log(f"fp32 {n} round trip {ms:.4f} ms, rel err {err:.3e}; device copy {copy_gbps:.0f} GB/s")
    for k, (tot, cnt, b) in sorted(prof.items(), key=lambda kv: -kv[1][0]):
        log(f"  {k:40s} avg {tot / max(cnt, 1) * 1e3:8.1f} us  {b / (tot / max(cnt, 1)) / 1e6:8.1f} GB/s"
            f"  x{cnt}")

    alg, comp = alg_bytes_roundtrip(n, 4)
    # dominant kernel = largest total time in the profiled pass
    dom = max(prof.items(), key=lambda kv: kv[1][0])
    dname, (dtot, dcnt, dbytes) = dom
    davg_ms = dtot / max(dcnt, 1)
    achieved = dbytes / (davg_ms * 1e-3) / 1e9

    extra = {}
    if not args.no_fp64:
        ms64, err64, _, _, _, unf64 = run_gpu(n, af.F64, max(5, args.steps // 2), 2, profile=False)
        alg64, _ = alg_bytes_roundtrip(n, 8)
        extra = {"fp64_ms_per_step": round(ms64, 4), "fp64_rel_err": err64,
                 "fp64_separate_z_y_ms": round(unf64, 4) if unf64 else None,
                 "fp64_pass_model_frac": round(alg64 / (ms64 * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        log(f"fp64 {n} round trip {ms64:.4f} ms, rel err {err64:.3e}")

    split, split_out = None, None
    if not args.no_split:
        log("split-precision matrix-core stages vs FP32 VALU (forward transforms):")
        dists = tuple(d for d in args.split_dists.split(",") if d)
        if not dists or any(d not in SPLIT_DISTS for d in dists) or "normal" not in dists:
            ap.error(f"--split-dists: a subset of {SPLIT_DISTS} that includes normal")
        split, split_out = run_split(n, max(5, args.steps // 2), dists)
        extra["split_precision"] = split
    if not args.no_wide:
        log("larger shapes:")
        extra["wide"] = run_wide(max(5, args.steps // 2))
    if not args.no_ops:
        log("operators (512^3 FP32), 256^3 round trips, differentiation (256^3 FP32):")
        extra["operators"] = run_ops(max(5, args.steps // 2))

    traffic, traffic_src = profiled_traffic(dname, n)
    out = {
        "metric": "512^3 r2c+c2r ms & %HBM peak; split-BF16 vs FP32 speedup at rel-err<=1e-5",
        "value": round(ms, 4),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded N(0,1) field resident in HBM)",
        "config": {
            "workload": f"{n[0]}^3 FP32 r2c+c2r round trip, out-of-place, 1x1 grid",
            "shape": list(n),
            "roundtrip_rel_l2": err,
            "pass_model_bytes": alg,
            "pass_model_frac_of_hbm_peak": round(alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "compulsory_frac_of_hbm_peak": round(comp / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "device_copy_GBps": round(copy_gbps, 1),
            "tile_tuning_xyz": tuning,
            "separate_z_y_passes_ms": round(unfused_ms, 4) if unfused_ms else None,
            "kernels": {k: {"avg_us": round(v[0] / max(v[1], 1) * 1e3, 2),
                            "alg_GBps": round(v[2] / (v[0] / max(v[1], 1) * 1e-3) / 1e9, 1)}
                        for k, v in prof.items()},
            **extra,
        },
        "roofline": {
            "kernel": dname,
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_source": traffic_src,
            "alg_bytes_per_launch": dbytes,
            "avg_launch_us": round(davg_ms * 1e3, 2),
        },
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        try:
            log("cpu baseline:")
            out["cpu_baseline"] = cpu_baseline(n, args.cpu_full)
            lib32 = out["cpu_baseline"]["library_torch_fft_round_trip"].get(f"{n[0]}^3_fp32_ms")
            if lib32:
                out["config"]["gpu_speedup_vs_cpu_library_fp32"] = round(lib32 / ms, 1)
        except Exception as e:  # the baseline is reported, never the measured path
            log(f"cpu baseline failed: {e!r}")
        if split:
            log("split-precision errors against the FP64 oracle:")
            check_split(n, split, split_out)
    split_out = None
    if split:
        b = split["bf16x2"]
        out["split_bf16_speedup_c2c"] = b["speedup_c2c_vs_fp32"]
        out["split_bf16_speedup_r2c"] = b["speedup_r2c_vs_fp32"]
        out["split_bf16_rel_err_max"] = b.get("rel_l2_max")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
