"""Memory-driven release of the intermediate spectra (the plans'
workspaces) in the autograd binding: the retry-once rule on allocation
failures (CPU, with stand-in callables) and the low-water / keep policy on
the GPU."""
import pytest


def test_retry_on_alloc_failure_releases_and_retries(monkeypatch):
    import accfft_amd as af
    from accfft_amd import autograd as ag
    released = []
    monkeypatch.setattr(ag, "release_all", lambda: released.append(1))
    monkeypatch.setattr(ag.torch.cuda, "empty_cache", lambda: None)
    calls = []

    def flaky():
        calls.append(1)
        if len(calls) == 1:
            raise af.AccfftError("device allocation failed", af.ERR_ALLOC)
        return "ok"
    assert ag._retry_on_oom(flaky) == "ok"
    assert len(calls) == 2 and released == [1]


def test_other_errors_are_not_retried(monkeypatch):
    import accfft_amd as af
    from accfft_amd import autograd as ag
    monkeypatch.setattr(ag, "release_all", lambda: pytest.fail("must not release"))
    calls = []

    def bad():
        calls.append(1)
        raise af.AccfftError("invalid argument", 1)
    with pytest.raises(af.AccfftError):
        ag._retry_on_oom(bad)
    assert len(calls) == 1


def test_second_failure_propagates(monkeypatch):
    import accfft_amd as af
    from accfft_amd import autograd as ag
    monkeypatch.setattr(ag, "release_all", lambda: None)
    monkeypatch.setattr(ag.torch.cuda, "empty_cache", lambda: None)

    def always():
        raise af.AccfftError("device allocation failed", af.ERR_ALLOC)
    with pytest.raises(af.AccfftError):
        ag._retry_on_oom(always)


def test_policy_arguments():
    import accfft_amd as af
    from accfft_amd import autograd as ag
    with pytest.raises(af.AccfftError):
        ag.set_memory_policy(low_water=-1)
    ag.set_memory_policy(low_water=None, keep=True)
    assert ag._POLICY == {"low_water": None, "keep": True}
    ag.set_memory_policy()
    assert ag._POLICY == {"low_water": 1 << 30, "keep": True}


@pytest.mark.gpu
def test_low_water_drops_spectra(dev):
    """With the low-water mark above the device's free memory, each call
    drops its plan's workspace; the results are the same as with the
    workspace kept."""
    import torch
    from accfft_amd import autograd as ag
    n = (64, 48, 40)
    S = ag.Spectral(n, torch.float64, dev)
    u = torch.randn(n, dtype=torch.float64, device=dev)
    try:
        ag.set_memory_policy(low_water=None, keep=True)
        ref = ag.inv_laplace(ag.laplace(u, S), S)
        torch.cuda.synchronize()
        assert S.plan.workspace_bytes() > 0
        free, total = torch.cuda.mem_get_info(dev)
        ag.set_memory_policy(low_water=total, keep=True)  # always "low"
        r0 = S.releases
        out = ag.inv_laplace(ag.laplace(u, S), S)
        torch.cuda.synchronize()
        assert S.plan.workspace_bytes() == 0 and S.releases >= r0 + 1
        assert torch.equal(out, ref)
        ag.set_memory_policy(low_water=None, keep=False)  # drop after every call
        out = ag.laplace(u, S)
        torch.cuda.synchronize()
        assert S.plan.workspace_bytes() == 0
    finally:
        ag.set_memory_policy()
    S.release()
