"""hipGraph replay of whole-run step loops (PAM_OPT_GRAPH): the replayed
forward run must give bit for bit the seismograms of the same launches issued
one by one, on the first call (capture), on later calls (replay of the cached
graph, with the CPML psi buffers' parity carried across an odd number of
steps), when batches of one shot share one graph through the staging buffer,
and after the model changes (the cached graph is dropped and re-captured)."""
import numpy as np
import pytest

from oracle.oracle import Oracle, rel_l2
from pseudoacousticmodeling_amd import _lib
from tests.problems import make_problem, make_engine, oracle_forward

pytestmark = pytest.mark.gpu


def _fwd(eng, nshots):
    return eng.forward(0, nshots).cpu().numpy()


@pytest.mark.parametrize("phys,abc", [(0, 0), (1, 1), (2, 1)])
def test_graph_replay_is_bit_identical(phys, abc):
    print(f"case graph phys={phys} abc={abc}", flush=True)
    pb = make_problem(phys, abc, nshots=3, seed=500 + 2 * phys + abc, nt=241)  # odd nt
    eng = make_engine(pb, max_shots=3)
    eng.set_option(_lib.PAM_OPT_GRAPH, 0)
    ref = _fwd(eng, 3)
    eng.set_option(_lib.PAM_OPT_GRAPH, 1)
    for call in range(3):  # capture, then replays
        got = _fwd(eng, 3)
        assert np.array_equal(got, ref), f"call {call}"
    # batches of one shot: one graph, three replays
    eng.set_option(_lib.PAM_OPT_BATCH, 1)
    got1 = _fwd(eng, 3)
    eng.set_option(_lib.PAM_OPT_GRAPH, 0)
    ref1 = _fwd(eng, 3)
    assert np.array_equal(got1, ref1)
    orc = oracle_forward(Oracle("f32"), pb, shot=2)["seis"]
    assert rel_l2(got1[2], orc) <= 1e-4


def test_graph_recaptured_after_model_change():
    pb = make_problem(1, 1, nshots=2, seed=510, nt=200)
    eng = make_engine(pb, max_shots=2)
    first = _fwd(eng, 2)
    _fwd(eng, 2)  # replay
    pb2 = dict(pb)
    pb2["vp"] = (pb["vp"] * 1.05).astype(np.float32)
    eng.set_model(pb2["vp"], pb2["eps"], pb2["dl"], pb2["th"])
    got = _fwd(eng, 2)
    eng.set_option(_lib.PAM_OPT_GRAPH, 0)
    ref = _fwd(eng, 2)
    assert np.array_equal(got, ref)
    assert not np.array_equal(got, first)
    orc = oracle_forward(Oracle("f32"), pb2, shot=1)["seis"]
    assert rel_l2(got[1], orc) <= 1e-4


def test_profiled_run_reports_kernel_time():
    pb = make_problem(0, 1, nshots=2, seed=520, nt=150)
    eng = make_engine(pb, max_shots=2)
    eng.set_option(_lib.PAM_OPT_LANES, 1)  # one batch of both shots
    eng.set_option(_lib.PAM_OPT_PROFILE, 1)
    a = _fwd(eng, 2)
    t = eng.timing()
    assert t["kernel_launches"] == pb["g"].nt and 0 < t["kernel_ms"] <= t["loop_ms"]
    eng.set_option(_lib.PAM_OPT_PROFILE, 0)
    b = _fwd(eng, 2)
    assert np.array_equal(a, b)
