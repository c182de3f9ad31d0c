"""A plan's r2c / c2r captured in a HIP graph (torch.cuda.CUDAGraph) and
replayed: every replay, with the program's own eager work and
synchronisations between replays, must equal the eager execution on the
same input bit for bit and match the FP64 oracle (src/accfft_gpu.cpp:444-623,
rel-L2 <= 1e-5 FP32).  The plan launches only kernels on its stream (no
memsets, no allocations once its workspaces exist); the persistent strided
tiles leave their tile queue zeroed at the end of every launch, so each
replay finds it as the capture did.  (2048, 40, 254) sends the x pass
through the queue; (48, 36, 40) uses one-shot tiles only.
"""
import numpy as np
import pytest

from conftest import rel_l2, rng

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(2048, 40, 254), (48, 36, 40)])
def test_graph_replays_match_eager_and_oracle(dev, orc, shape):
    import torch
    import accfft_amd as af
    print(f"graph capture shape={shape}", flush=True)
    nh = shape[2] // 2 + 1
    x = torch.empty(shape, dtype=torch.float32, device=dev)
    h = torch.empty((shape[0], shape[1], nh), dtype=torch.complex64, device=dev)
    h2 = torch.empty_like(h)
    y = torch.empty_like(x)
    plan = af.Plan("r2c", shape, x, h, af.F32, af.ACCFFT_ESTIMATE)

    def step():
        plan.execute_r2c(x, h)
        h2.copy_(h)  # c2r overwrites its input
        plan.execute_c2r(h2, y)

    inputs = [rng(700 + i).standard_normal(shape).astype(np.float32) for i in range(3)]
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        # eager warm-up on the capture stream: binds the plan to it
        x.copy_(torch.from_numpy(inputs[0]).to(dev))
        step()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        step()
    torch.cuda.synchronize()
    for i, inp in enumerate(inputs):
        xin = torch.from_numpy(inp).to(dev)
        with torch.cuda.stream(s):
            x.copy_(xin)
            g.replay()
        torch.cuda.synchronize()
        hg, yg = h.clone(), y.clone()
        with torch.cuda.stream(s):
            x.copy_(xin)
            step()
        torch.cuda.synchronize()
        assert torch.equal(hg, h), i
        assert torch.equal(yg, y), i
        ref = orc.r2c(inp.astype(np.float64))
        assert rel_l2(hg.cpu().numpy(), ref) < 1e-5, i
        assert rel_l2(yg.cpu().numpy() / np.prod(shape), inp) < 1e-5, i
    del g
    plan.destroy()
