"""Shared fixtures.  Tests marked ``gpu`` need an MI355X and the built
libaccfft_amd.so; everything else runs on the CPU."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU and libaccfft_amd.so")


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import accfft_amd
    accfft_amd.lib()  # fails loudly when the HIP library is missing
    return torch.device("cuda:0")


def rel_l2(a, b) -> float:
    a = np.asarray(a)
    b = np.asarray(b)
    nb = np.linalg.norm(b.ravel())
    d = np.linalg.norm((a - b).ravel())
    return float(d / nb) if nb > 0 else float(d)


def rng(seed: int):
    return np.random.default_rng(seed)


def pytest_runtest_logstart(nodeid, location):
    """GPU runs print a marker before and after each test on stdout, so that
    the tail of a run that ends abruptly names the test that was running."""
    if os.environ.get("ACCFFT_AMD_TEST_MARKERS"):
        print(f"\n[START] {nodeid}", flush=True)


def pytest_runtest_logreport(report):
    if os.environ.get("ACCFFT_AMD_TEST_MARKERS") and report.when == "call":
        print(f"[END] {report.nodeid} {report.outcome}", flush=True)
