"""The pencil share across processes: one process per rank of a p0 x p1
grid (all on the one GPU), the exchange hook an all_to_all_single over
torch.distributed process groups for the row and column communicators
(gloo, host-staged), each rank's block against the oracle's global
transform (tests/dist_pencil_worker.py)."""
import os
import socket
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _port():
    with socket.socket() as s:
        s.bind(("127.0.0.1", 0))
        return s.getsockname()[1]


@pytest.mark.parametrize("N,grid", [((9, 10, 7), (2, 2)), ((8, 16, 20), (2, 2)), ((6, 9, 10), (3, 1))])
def test_pencil_share_over_processes(N, grid):
    nproc = grid[0] * grid[1]
    print(f"pencil share {N} grid {grid}: {nproc} processes", flush=True)
    env = dict(os.environ, PENCIL_N=",".join(map(str, N)), PENCIL_GRID=",".join(map(str, grid)),
               MASTER_ADDR="127.0.0.1", OMP_NUM_THREADS="1")
    cmd = [sys.executable, "-m", "torch.distributed.run", "--nnodes=1",
           f"--nproc-per-node={nproc}", "--master-addr=127.0.0.1", f"--master-port={_port()}",
           os.path.join(HERE, "dist_pencil_worker.py")]
    r = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=240)
    print(r.stdout[-3000:])
    assert r.returncode == 0, r.stderr[-3000:]
    assert r.stdout.count(" ok") == nproc
