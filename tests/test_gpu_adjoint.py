"""GPU parity of the adjoint passes (RTM image + illumination, FWI gradient)
against the oracle's on-the-fly restatement of src/Migration.py:963-1111,
through the C ABI."""
import numpy as np
import pytest

from oracle.oracle import Oracle, rel_l2, rtm_onthefly
from tests.problems import make_problem, make_engine, oracle_forward

pytestmark = pytest.mark.gpu

# float32 GPU vs float32 restatement: relative L2 of image / gradient
TOL = 1e-4


def _data(pb, nshots, seed):
    """Observed data for the adjoint: the oracle's seismograms of a perturbed
    model (the realistic amplitude range of a residual / muted record)."""
    rng = np.random.default_rng(seed)
    pert = dict(pb)
    pert["vp"] = (pb["vp"] * (1.0 + 0.05 * rng.standard_normal(pb["vp"].shape))).astype(np.float32)
    orc = Oracle("f64")
    return np.stack([oracle_forward(orc, pert, shot=s)["seis"] for s in range(nshots)]).astype(np.float32)


def _oracle_adjoint(pb, data, fwi, multiparameter=False, kind="f32", **storage):
    orc = Oracle(kind)
    g = pb["g"]
    model = tuple(orc.expand(m, g.N) if m is not None else None for m in (pb["vp"], pb["eps"], pb["dl"], pb["th"]))
    A = orc.cerjan(g.N) if pb["abc"] == 0 else None
    cpml = (pb["d0"], pb["fp"]) if pb["abc"] == 1 else None
    shots = list(zip(pb["sz"].tolist(), pb["sx"].tolist()))
    interior = (pb["vp"], pb["eps"], pb["dl"], pb["th"])
    return rtm_onthefly(orc, pb["phys"], pb["abc"], g, model, pb["src"], shots, pb["rz"], pb["rx"], data, A=A,
                        cpml=cpml, fwi=fwi, multiparameter=multiparameter, model_interior=interior, **storage)


@pytest.mark.parametrize("phys,abc", [(p, a) for p in (0, 1, 2) for a in (0, 1)])
def test_rtm_matches_oracle(phys, abc):
    print(f"case rtm phys={phys} abc={abc}", flush=True)
    pb = make_problem(phys, abc, nshots=2, seed=70 + 2 * phys + abc, nt=240)
    data = _data(pb, 2, seed=7)
    ref = _oracle_adjoint(pb, data, fwi=False)
    eng = make_engine(pb, max_shots=2)
    image, ilum = eng.rtm(data)
    image = image.cpu().numpy()
    ilum = ilum.cpu().numpy()
    e_img = rel_l2(image, ref["image"])
    e_ilum = rel_l2(ilum, ref["ilum"])
    ref64 = _oracle_adjoint(pb, data, fwi=False, kind="f64")
    print(f"  image rel_l2 {e_img:.3e}, ilum rel_l2 {e_ilum:.3e}; float64 oracle: image "
          f"{rel_l2(image, ref64['image']):.3e}, ilum {rel_l2(ilum, ref64['ilum']):.3e}", flush=True)
    assert np.isfinite(image).all() and np.abs(image).max() > 0
    assert e_ilum <= TOL
    assert e_img <= TOL


@pytest.mark.parametrize("phys,abc", [(p, a) for p in (0, 1, 2) for a in (0, 1)])
def test_gradient_matches_oracle(phys, abc):
    print(f"case gradient phys={phys} abc={abc}", flush=True)
    pb = make_problem(phys, abc, nshots=2, seed=80 + 2 * phys + abc, nt=240)
    data = _data(pb, 2, seed=8)
    ref = _oracle_adjoint(pb, data, fwi=True)
    eng = make_engine(pb, max_shots=2)
    out = eng.gradient(data)
    g = out["g_m"].cpu().numpy()
    err = rel_l2(g, ref["image"])
    e64 = rel_l2(g, _oracle_adjoint(pb, data, fwi=True, kind="f64")["image"])
    print(f"  gradient rel_l2 {err:.3e}; float64 oracle {e64:.3e}", flush=True)
    assert np.isfinite(g).all() and np.abs(g).max() > 0
    assert err <= TOL


def test_rtm_batched_equals_one_shot_at_a_time():
    pb = make_problem(0, 1, nshots=3, seed=90, nt=200)
    data = _data(pb, 3, seed=9)
    eng = make_engine(pb, max_shots=3)
    img3, il3 = eng.rtm(data)
    from pseudoacousticmodeling_amd import _lib

    eng.set_option(_lib.PAM_OPT_BATCH, 1)
    img1, il1 = eng.rtm(data)
    assert rel_l2(img1.cpu().numpy(), img3.cpu().numpy()) < 1e-6
    assert rel_l2(il1.cpu().numpy(), il3.cpu().numpy()) < 1e-6


@pytest.mark.parametrize("phys,abc", [(1, 0), (1, 1), (2, 0), (2, 1)])
def test_multiparameter_partials_match_oracle(phys, abc):
    """eps / delta (/ theta) partials of the FWI gradient, src/Migration.py:1053-1064 with
    src/CudaKernels.py:1174-1379, normalised by the illumination."""
    print(f"case partials phys={phys} abc={abc}", flush=True)
    pb = make_problem(phys, abc, nshots=2, seed=95 + 2 * phys + abc, nt=240)
    data = _data(pb, 2, seed=11)
    ref = _oracle_adjoint(pb, data, fwi=True, multiparameter=True)
    eng = make_engine(pb, max_shots=2)
    out = eng.gradient(data, multiparameter=True)
    names = [("g_eps", "g_eps"), ("g_delta", "g_delta")] + ([("g_theta", "g_theta")] if phys == 2 else [])
    assert rel_l2(out["g_m"].cpu().numpy(), ref["image"]) <= TOL
    ref64 = _oracle_adjoint(pb, data, fwi=True, multiparameter=True, kind="f64")
    for gk, rk in names:
        g = out[gk].cpu().numpy()
        err = rel_l2(g, ref[rk])
        print(f"  {gk} rel_l2 {err:.3e}; float64 oracle {rel_l2(g, ref64[rk]):.3e}", flush=True)
        assert np.isfinite(g).all() and np.abs(g).max() > 0
        assert err <= TOL
