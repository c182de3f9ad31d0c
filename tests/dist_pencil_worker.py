"""One rank of a pencil-share r2c / c2r over torch.distributed (run by
tests/test_gpu_pencil_dist.py under torch.distributed.run, one process per
rank, all on the one GPU).

The plan's exchange_v hook is the integration a caller writes: the row and
column communicators are process groups of the ranks sharing c0 / c1 (the
reference's row_comm / col_comm, src/accfft_gpu.cpp:86-87), and each
exchange is an all_to_all_single with per-peer split sizes.  The gloo
backend takes host tensors, so the hook stages the blocks through host
memory with synchronous copies (with the nccl backend = RCCL the device
buffers would be passed directly).  Exits non-zero when the rank's block
differs from the oracle's global transform."""
import ctypes
import os
import sys

import numpy as np
import torch
import torch.distributed as dist

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, HERE)


def _hip():
    with open("/proc/self/maps") as f:
        for line in f:
            if "libamdhip64.so" in line:
                hip = ctypes.CDLL(line.split()[-1])
                hip.hipMemcpyAsync.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                               ctypes.c_int, ctypes.c_void_p]
                hip.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
                return hip
    raise RuntimeError("libamdhip64 not loaded")


def main():
    import accfft_amd as af
    import oracle as orc
    N = tuple(int(v) for v in os.environ["PENCIL_N"].split(","))
    P0, P1 = (int(v) for v in os.environ["PENCIL_GRID"].split(","))
    dist.init_process_group("gloo")
    rank, world = dist.get_rank(), dist.get_world_size()
    assert world == P0 * P1
    c0, c1 = divmod(rank, P1)  # row-major coordinates, as MPI_Cart_create
    rows = [dist.new_group([a * P1 + b for b in range(P1)]) for a in range(P0)]
    cols = [dist.new_group([a * P1 + b for a in range(P0)]) for b in range(P1)]
    hip = _hip()
    calls = []

    def hook(comm, send, sb, so, recv, rb, ro, direction, stream):
        calls.append((comm, direction))
        if hip.hipStreamSynchronize(stream) != 0:
            return 5
        sh = torch.empty(sum(sb), dtype=torch.uint8)
        rh = torch.empty(sum(rb), dtype=torch.uint8)
        # device -> host on the plan's stream, waited for before gloo reads it
        for q in range(len(sb)):
            if sb[q] and hip.hipMemcpyAsync(sh.data_ptr() + sum(sb[:q]), send + so[q], sb[q], 2,
                                            stream) != 0:
                return 5
        if hip.hipStreamSynchronize(stream) != 0:
            return 5
        dist.all_to_all_single(rh, sh, output_split_sizes=list(rb), input_split_sizes=list(sb),
                               group=rows[c0] if comm == 0 else cols[c1])
        # host -> device on the plan's stream, finished before rh goes away
        for q in range(len(rb)):
            if rb[q] and hip.hipMemcpyAsync(recv + ro[q], rh.data_ptr() + sum(rb[:q]), rb[q], 1,
                                            stream) != 0:
                return 5
        return 0 if hip.hipStreamSynchronize(stream) == 0 else 5

    g = np.random.default_rng(20261202).standard_normal(N)
    G = orc.r2c(g)
    torch.cuda.set_device(0)
    p = af.Plan("r2c_pencil", N, None, None, af.F64, af.ACCFFT_ESTIMATE, grid=(P0, P1),
                coords=(c0, c1))
    p.set_exchange_v(hook)
    n0, n1, _ = p.isize
    xs, y1s, _ = p.istart
    x = torch.tensor(np.ascontiguousarray(g[xs:xs + n0, y1s:y1s + n1]), device="cuda:0")
    out = torch.empty(p.osize, dtype=torch.complex128, device="cuda:0")
    timer = [0.0] * 5
    p.execute_r2c(x, out, timer=timer)
    torch.cuda.synchronize()
    _, ys, zs = p.ostart
    ref = G[:, ys:ys + p.osize[1], zs:zs + p.osize[2]]
    err = 0.0
    if ref.size:
        err = np.linalg.norm(out.cpu().numpy() - ref) / np.linalg.norm(ref)
    back = torch.empty_like(x)
    p.execute_c2r(out, back)
    torch.cuda.synchronize()
    err2 = 0.0
    if back.numel():
        b = back.cpu().numpy() / np.prod(N)
        err2 = np.linalg.norm(b - g[xs:xs + n0, y1s:y1s + n1]) / np.linalg.norm(b)
    p.destroy()
    ok = err < 1e-12 and err2 < 1e-12 and len(calls) == 2 * ((P0 > 1) + (P1 > 1))
    print(f"rank {rank} ({c0},{c1}) forward rel-L2 {err:.2e} round trip {err2:.2e} "
          f"exchanges {calls} comm time {timer[2]:.2e} s {'ok' if ok else 'FAIL'}", flush=True)
    dist.barrier()
    dist.destroy_process_group()
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
