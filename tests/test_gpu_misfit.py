"""The misfit X = sum_shots 1/2 sum (dobs - dcal)^2 (src/Inversion.py:13-53):
pam_misfit's residual is the float32 difference bit for bit, its float64
reduction agrees with a float64 numpy sum of the oracle's residual, and two
calls give the same bits (the two-pass reduction has a fixed order)."""
import numpy as np
import pytest
import torch

from oracle.oracle import Oracle
from tests.problems import make_problem, make_engine, oracle_forward

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("phys,abc", [(0, 0), (1, 1)])
def test_misfit_matches_float64_sum_of_oracle_residual(phys, abc):
    print(f"case misfit phys={phys} abc={abc}", flush=True)
    pb = make_problem(phys, abc, nshots=3, seed=901 + phys, nt=240)
    eng = make_engine(pb, max_shots=3)
    dcal = eng.forward(0, 3)
    orc = Oracle("f32")
    # observed data: the oracle's seismograms of a perturbed model
    pb2 = dict(pb)
    pb2["vp"] = (pb["vp"] * np.float32(1.03)).astype(np.float32)
    dobs = np.stack([oracle_forward(orc, pb2, shot=s)["seis"] for s in range(3)]).astype(np.float32)
    X, res = eng.misfit(dobs, dcal)
    dc = dcal.cpu().numpy()
    r_ref = dobs - dc  # float32 difference, as the kernel forms it
    assert np.array_equal(res.cpu().numpy(), r_ref)
    X64 = float(np.sum(0.5 * r_ref.astype(np.float64) ** 2))
    print(f"misfit gpu {X!r} float64 numpy {X64!r} rel {abs(X - X64) / X64:.3e}", flush=True)
    assert X64 > 0 and abs(X - X64) <= 1e-12 * X64
    X2, _ = eng.misfit(dobs, dcal)
    assert X2 == X  # bit for bit
    eng.close()


def test_misfit_of_large_gathers_is_reproducible():
    """More samples than blocks of the first pass (every block strides)."""
    pb = make_problem(0, 0, nshots=1, nt=120, seed=911)
    eng = make_engine(pb)
    gen = torch.Generator(device="cuda").manual_seed(5)
    n = 5_000_000
    a = torch.randn(n, device="cuda", generator=gen)
    b = torch.randn(n, device="cuda", generator=gen)
    # reinterpret as gathers of this grid's row length (nshots rows of nt_data * nrec)
    row = eng.nt_data * eng.nrec
    ns = n // row
    a = a[: ns * row].view(ns, eng.nt_data, eng.nrec)
    b = b[: ns * row].view(ns, eng.nt_data, eng.nrec)
    X1, r1 = eng.misfit(a, b)
    X2, r2 = eng.misfit(a, b)
    assert X1 == X2
    r = (a - b).cpu().numpy()
    X64 = float(np.sum(0.5 * r.astype(np.float64) ** 2))
    assert abs(X1 - X64) <= 1e-11 * X64
    eng.close()
