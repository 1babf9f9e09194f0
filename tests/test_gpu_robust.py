"""Failure handling forced from the host: a time step far above the stability
limit makes the wavefield overflow to Inf / NaN (ordinary float arithmetic,
every kernel exits normally) and the whole-run entry points then return
PAM_E_NONFINITE; a stable run of the same engine is not flagged.  Groups
re-created over the same engines (graphs on) replay nothing of a destroyed
group."""
import numpy as np
import pytest

from pseudoacousticmodeling_amd import _lib
from tests.problems import make_problem, make_engine

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("phys,abc", [(0, 0), (1, 1)])
def test_unstable_run_reports_nonfinite(phys, abc):
    print(f"case nonfinite phys={phys} abc={abc}", flush=True)
    # vp up to 3000 m/s, dx 10 m: the 8th-order limit is about 1.1 ms; 8 ms blows up
    pb = make_problem(phys, abc, nt=400, dt=0.008, nshots=2, seed=31)
    eng = make_engine(pb, max_shots=2)
    with pytest.raises(_lib.PamError, match="PAM_E_NONFINITE"):
        eng.forward(0, 2)
    res = np.zeros((2, eng.nt_data, eng.nrec), np.float32)
    with pytest.raises(_lib.PamError, match="PAM_E_NONFINITE"):
        eng.rtm(res + 1e-3)
    eng.close()
    ok = make_engine(make_problem(phys, abc, nt=400, nshots=2, seed=31), max_shots=2)
    seis = ok.forward(0, 2)
    assert np.isfinite(seis.cpu().numpy()).all()
    ok.close()


def test_group_recreated_over_the_same_engines():
    """A group's step-loop graphs read its device tables: destroying the group
    destroys them, and a new group over the same engines (same epochs, same
    output buffers) builds its own."""
    from pseudoacousticmodeling_amd.engine import Group

    shots = [2, 1]
    pbs = [make_problem(1, 1, nz=nz, nx=nx, nshots=k, seed=720 + i, nt=150)
           for i, ((nz, nx), k) in enumerate(zip([(40, 56), (32, 76)], shots))]
    engs = [make_engine(pb, max_shots=k) for pb, k in zip(pbs, shots)]
    for e in engs:
        e.set_option(_lib.PAM_OPT_TILE, 0)
        e.set_option(_lib.PAM_OPT_GRAPH, 1)
    refs = [e.forward(0, k).cpu().numpy() for e, k in zip(engs, shots)]
    outs = None
    for rep in range(3):
        grp = Group(engs)
        outs = grp.forward(shots, outs)  # the same output buffers every time
        for i, (o, r) in enumerate(zip(outs, refs)):
            assert np.array_equal(o.cpu().numpy(), r), f"member {i} differs (group {rep})"
        grp.close()
