"""pam_misfit_noalloc, the allocation-free misfit: called through the C ABI it
gives pam_misfit's residual and float64 misfit bit for bit, on two calls in a
row (the second with fewer shots and fewer samples, so that stale partials of
the first call in the shared scratch array would show), with and without a
residual output.  Its size checks run before any device call."""
import ctypes

import numpy as np
import pytest
import torch

from pseudoacousticmodeling_amd import _lib
from tests.problems import make_problem, make_engine


def _noalloc(lib, dobs, dcal, residual, out):
    ns, nt, nr = dobs.shape
    stream = torch.cuda.current_stream().cuda_stream
    rc = lib.pam_misfit_noalloc(ns, nt, nr, ctypes.c_void_p(dobs.data_ptr()), ctypes.c_void_p(dcal.data_ptr()),
                                ctypes.c_void_p(residual.data_ptr()) if residual is not None else None,
                                ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(stream))
    _lib.check(None, rc)
    torch.cuda.current_stream().synchronize()
    return float(out.cpu().item())


def test_rejects_sizes_beyond_the_maximum_before_any_device_call():
    lib = _lib.load()
    big = 1 << 14  # 2^14 * 2^14 * 2^13 = 2^41 samples > 2^40
    rc = lib.pam_misfit_noalloc(big, big, 1 << 13, ctypes.c_void_p(16), ctypes.c_void_p(16), None,
                                ctypes.c_void_p(16), None)
    assert rc == -1
    assert b"exceed the maximum" in lib.pam_last_error(None)
    assert lib.pam_misfit_noalloc(0, 10, 10, ctypes.c_void_p(16), ctypes.c_void_p(16), None,
                                  ctypes.c_void_p(16), None) == -1
    assert lib.pam_misfit_noalloc(1, 10, 10, None, ctypes.c_void_p(16), None, ctypes.c_void_p(16), None) == -1


@pytest.mark.gpu
def test_noalloc_misfit_equals_pam_misfit_on_two_calls():
    lib = _lib.load()
    pb = make_problem(1, 1, nshots=4, seed=977, nt=220)
    eng = make_engine(pb, max_shots=4)
    gen = torch.Generator(device="cuda").manual_seed(11)
    out = torch.empty(1, dtype=torch.float64, device="cuda")
    # first call: 4 shots of seismograms against noisy copies; second: 2 shots
    # (fewer samples and fewer partial sums than the first call left behind)
    dcal = eng.forward(0, 4)
    dobs = dcal + 1e-3 * torch.randn(dcal.shape, device="cuda", generator=gen)
    for call, ns in enumerate([4, 2]):
        a, b = dobs[:ns].contiguous(), dcal[:ns].contiguous()
        print(f"case noalloc call={call} shape={tuple(a.shape)}", flush=True)
        X_ref, r_ref = eng.misfit(a, b)
        res = torch.empty_like(a)
        X = _noalloc(lib, a, b, res, out)
        assert torch.equal(res, r_ref)
        assert X == X_ref, (X, X_ref)
        X64 = float(np.sum(0.5 * (a - b).cpu().numpy().astype(np.float64) ** 2))
        print(f"noalloc {X!r} pam_misfit {X_ref!r} float64 numpy {X64!r}", flush=True)
        assert abs(X - X64) <= 1e-12 * X64
        assert _noalloc(lib, a, b, None, out) == X
    eng.close()
