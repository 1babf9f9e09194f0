"""The temporal-blocking variant (k_tblock, PAM_OPT_TILE 12): 4 steps per
launch on an LDS-staged tile with a 16-cell halo, for grids too small to fill
the device (Cerjan boundary).  Every cell goes through the 16-row tile's
arithmetic, so seismograms recorded densely (whole rows next to
the rim, on tile boundaries and in the interior) must be variant 0's bit for
bit, for step counts that are and are not multiples of the launch's steps,
sources in the absorbing layer and the interior, and the Cerjan taper on."""
import numpy as np
import pytest

from oracle.oracle import Oracle, rel_l2
from tests.problems import make_engine, make_problem, oracle_forward

pytestmark = pytest.mark.gpu

TBLOCK = 12  # the variant index (pam_variants.cuh)
CERJAN = 0.0045


def _problem(phys, nt, itlag, nshots, seed):
    pb = make_problem(phys, 0, nz=61, nx=197, N=12, nshots=nshots, seed=seed, nt=nt, itlag=itlag)
    g = pb["g"]
    nzp, nxp = g.nz + 2 * g.N, g.nx + 2 * g.N
    rows = [4, 5, 15, 16, 31, 32, g.N + 7, nzp - 5]
    cols = np.arange(4, nxp - 4, dtype=np.int32)
    pb["rz"] = np.concatenate([np.full(len(cols), r, dtype=np.int32) for r in rows])
    pb["rx"] = np.concatenate([cols for _ in rows]).astype(np.int32)
    # sources in the absorbing layer, on tile boundaries and in the interior
    sz = [5, 16, g.N + 20, nzp - 6, 47, 63, 30]
    sx = [6, 63, 64, nxp - 7, 128, 100, 190]
    pb["sz"] = np.array(sz[:nshots], dtype=np.int32)
    pb["sx"] = np.array(sx[:nshots], dtype=np.int32)
    return pb


@pytest.mark.parametrize("phys", [0, 1, 2])
@pytest.mark.parametrize("nt,itlag", [(123, 9), (37, 30), (40, 17), (16, 3), (5, 2)])
def test_temporal_blocking_is_bit_identical(phys, nt, itlag):
    from pseudoacousticmodeling_amd import _lib

    ns = 7
    pb = _problem(phys, nt, itlag, ns, seed=400 + phys)
    outs = {}
    for t in (0, TBLOCK):
        print(f"case tblock phys={phys} nt={nt} itlag={itlag} variant={t}", flush=True)
        eng = make_engine(pb, max_shots=ns, cerjan_factor=CERJAN)
        eng.set_option(_lib.PAM_OPT_TILE, t)
        first = eng.forward(0, ns).cpu().numpy()
        again = eng.forward(0, ns).cpu().numpy()  # the captured graph replayed
        eng.set_option(_lib.PAM_OPT_GRAPH, 0)
        direct = eng.forward(0, 3).cpu().numpy()  # launched one by one, a smaller batch
        outs[t] = (first, again, direct)
        eng.close()
    ref = outs[0]
    assert np.isfinite(ref[0]).all() and np.abs(ref[0]).max() > 0
    for i, name in enumerate(("first run", "graph replay", "direct launches")):
        assert np.array_equal(outs[TBLOCK][i], ref[i]), f"{name} differs from variant 0"


@pytest.mark.parametrize("phys", [0, 2])
def test_temporal_blocking_matches_oracle(phys):
    """The seismogram against the float32 oracle (relative L2 <= 1e-4)."""
    from pseudoacousticmodeling_amd import _lib

    pb = make_problem(phys, 0, nz=45, nx=150, nshots=2, seed=420 + phys, nt=203)
    eng = make_engine(pb, max_shots=2, cerjan_factor=CERJAN)
    eng.set_option(_lib.PAM_OPT_TILE, TBLOCK)
    seis = eng.forward(0, 2).cpu().numpy()
    eng.close()
    for s in range(2):
        ref = oracle_forward(Oracle("f32"), pb, shot=s, cerjan_factor=CERJAN)["seis"]
        err = rel_l2(seis[s], ref)
        ref64 = oracle_forward(Oracle("f64"), pb, shot=s, cerjan_factor=CERJAN)["seis"]
        print(f"  phys={phys} shot {s}: rel_l2 vs f32 oracle {err:.3e}, vs f64 {rel_l2(seis[s], ref64):.3e}",
              flush=True)
        assert err <= 1e-4
