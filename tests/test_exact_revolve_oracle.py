"""The oracle's restatement of the revolve-checkpointed exact adjoint
(oracle/exact_adjoint.py revolve_states): with float32 slots its states are
the full forward pass's, whatever the slot count (the treeverse schedule
recomputes every state it hands out exactly); with float16 slots only the
restarts are rounded, so the gradient moves by about the float16 rounding and
no more."""
import numpy as np
import pytest

from oracle import exact_adjoint as EA
from tests.test_gpu_exact_adjoint import _problem, _setup


def _st(abc, nt=60):
    pb, _ = _problem(abc, 70 + abc, nshots=1, nt=nt)
    N = pb["g"].N
    return _setup(pb, 0), np.pad(pb["vp"].astype(np.float64), N, mode="edge")


@pytest.mark.parametrize("abc", [0, 1])
@pytest.mark.parametrize("slots", [1, 3, 7, 60])
def test_float32_slots_give_the_forward_states(abc, slots):
    st, vp = _st(abc)
    ref = EA.forward(st, vp, keep=True)["states"]
    states, prevs = EA.revolve_states(st, vp, slots, "f32")
    for k in range(st.nt):
        for name, v in ref[k].items():
            assert np.array_equal(states[k][name], v), (k, name)
        if abc == 1 and k >= 1:
            for name in ("px", "pz", "zx", "zz"):
                assert np.array_equal(prevs[k][name], ref[k - 1][name]), (k, name)


@pytest.mark.parametrize("abc", [0, 1])
def test_float16_slots_move_the_gradient_by_about_the_rounding(abc):
    st, vp = _st(abc)
    rng = np.random.default_rng(5)
    r = rng.standard_normal((st.nt - st.itlag, len(st.rz)))
    g = EA.adjoint(st, vp, r)
    errs = []
    for slots in (3, 12):
        states, prevs = EA.revolve_states(st, vp, slots, "f16")
        gq = EA.adjoint(st, vp, r, states=states, prev_states=prevs)
        errs.append(np.linalg.norm(gq - g) / np.linalg.norm(g))
    print("relative change with float16 slots (3, 12 slots):", ["%.1e" % e for e in errs])
    # changed (the restarts are rounded) but within a few float16 roundings
    assert all(1e-7 < e < 2e-2 for e in errs)
