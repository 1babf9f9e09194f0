"""GPU exact adjoint (pam_forward_adjoint) and the torch.autograd.Function over
it: parity with the float64 oracle adjoint (oracle/exact_adjoint.py, itself
pinned by dot-product and finite-difference tests), central finite
differences of the GPU's own forward misfit, and autograd plumbing."""
import numpy as np
import pytest
import torch

from oracle import exact_adjoint as EA
from oracle.oracle import Oracle, rel_l2
from pseudoacousticmodeling_amd import _lib
from tests.problems import make_engine, make_problem, smooth_field

pytestmark = pytest.mark.gpu

# float32 GPU gradient vs the float64 oracle (relative L2)
TOL = 1e-4


def _setup(pb, shot):
    g = pb["g"]
    return EA.Setup(nzp=g.nzp, nxp=g.nxp, N=g.N, dz=g.dz, dx=g.dx, dt=g.dt, nt=g.nt, itlag=g.itlag,
                    abc=pb["abc"], src=np.asarray(pb["src"], dtype=np.float64), sz=int(pb["sz"][shot]),
                    sx=int(pb["sx"][shot]), rz=np.asarray(pb["rz"]), rx=np.asarray(pb["rx"]),
                    A=np.asarray(Oracle("f64").cerjan(g.N, 0.0), dtype=np.float64), d0=pb["d0"], fpico=pb["fp"])


def _problem(abc, seed, nshots=2, nt=260):
    pb = make_problem(0, abc, nz=40, nx=56, N=12, nt=nt, itlag=30, nshots=nshots, seed=seed)
    rng = np.random.default_rng(seed + 1000)
    true = dict(pb)
    true["vp"] = (pb["vp"] * (1.0 + 0.03 * rng.standard_normal(pb["vp"].shape))).astype(np.float32)
    return pb, true


def _oracle_grad(pb, res):
    N = pb["g"].N
    vp = np.pad(pb["vp"].astype(np.float64), N, mode="edge")
    g = np.zeros_like(vp)
    for s in range(len(pb["sz"])):
        g += EA.adjoint(_setup(pb, s), vp, np.asarray(res[s], dtype=np.float64))
    return EA.expand_T(g, N), g


@pytest.mark.parametrize("abc,storage", [(0, _lib.PAM_STORE_ALL_F32), (1, _lib.PAM_STORE_ALL_F32),
                                         (0, _lib.PAM_STORE_CHECKPOINT), (1, _lib.PAM_STORE_CHECKPOINT)])
def test_exact_adjoint_matches_float64_oracle(abc, storage):
    print(f"case exact adjoint abc={abc} storage={storage}", flush=True)
    pb, true = _problem(abc, 30 + abc)
    eng = make_engine(true, max_shots=2)
    dobs = eng.forward(0, 2)
    eng.set_model(pb["vp"])
    eng.set_option(_lib.PAM_OPT_STORAGE, storage)
    if storage == _lib.PAM_STORE_CHECKPOINT:
        eng.set_option(_lib.PAM_OPT_CKPT_STEP, 23)
    res = eng.forward(0, 2) - dobs
    g = eng.forward_adjoint(res).cpu().numpy()
    gpad = eng.forward_adjoint(res, layout=_lib.PAM_PADDED).cpu().numpy()
    ref, ref_pad = _oracle_grad(pb, res.cpu().numpy())
    e = rel_l2(g, ref)
    e_pad = rel_l2(gpad, ref_pad)
    print(f"  interior rel_l2 {e:.2e}  padded rel_l2 {e_pad:.2e}", flush=True)
    assert np.isfinite(g).all() and np.abs(g).max() > 0
    assert e <= TOL and e_pad <= TOL
    np.testing.assert_allclose(EA.expand_T(gpad, pb["g"].N), g, rtol=1e-4, atol=1e-6 * np.abs(g).max())


def test_checkpointed_state_gives_the_same_gradient():
    pb, true = _problem(1, 40)
    eng = make_engine(true, max_shots=2)
    dobs = eng.forward(0, 2)
    eng.set_model(pb["vp"])
    res = eng.forward(0, 2) - dobs
    g_all = eng.forward_adjoint(res).cpu().numpy()
    eng.set_option(_lib.PAM_OPT_STORAGE, _lib.PAM_STORE_CHECKPOINT)
    for step in (0, 7, 50):
        eng.set_option(_lib.PAM_OPT_CKPT_STEP, step)
        g_ck = eng.forward_adjoint(res).cpu().numpy()
        # the recomputed states are bit-identical, so is the gradient
        np.testing.assert_array_equal(g_ck, g_all)


@pytest.mark.parametrize("abc", [0, 1])
def test_revolve_float32_slots_give_the_same_gradient(abc):
    pb, true = _problem(abc, 42 + abc)
    eng = make_engine(true, max_shots=2)
    dobs = eng.forward(0, 2)
    eng.set_model(pb["vp"])
    res = eng.forward(0, 2) - dobs
    g_all = eng.forward_adjoint(res).cpu().numpy()
    eng.set_option(_lib.PAM_OPT_STORAGE, _lib.PAM_STORE_REVOLVE_F32)
    for slots in (1, 7, 50):
        print(f"case exact revolve f32 abc={abc} slots={slots}", flush=True)
        eng.set_option(_lib.PAM_OPT_SNAPSHOTS, slots)
        np.testing.assert_array_equal(eng.forward_adjoint(res).cpu().numpy(), g_all)


@pytest.mark.parametrize("abc", [0, 1])
@pytest.mark.parametrize("slots", [7, 50])
def test_revolve_float16_slots_match_the_oracle_restatement(abc, slots):
    """float16 slots: the GPU gradient against the float64 oracle adjoint of the
    same treeverse sweep from float16-rounded restarts (oracle
    revolve_states).  Which way each value rounds to float16 depends on the
    last bits of the state (the oracle's float64 states and the GPU's float32
    ones differ there): a 1e-6 relative perturbation of the states before the
    rounding already moves the restated gradient by about half of the
    rounding's whole effect.  So the GPU is held to the statistics of the
    restatement, not to its bits: its distance to the exact float64 gradient
    is within a factor 3 of the restatement's (the slots are rounded, and by
    the same amount), and its distance to the restatement is no larger than
    twice the restatement's distance to the exact gradient."""
    print(f"case exact revolve f16 abc={abc} slots={slots}", flush=True)
    pb, true = _problem(abc, 44 + abc)
    eng = make_engine(true, max_shots=2)
    dobs = eng.forward(0, 2)
    eng.set_model(pb["vp"])
    res = eng.forward(0, 2) - dobs
    eng.set_option(_lib.PAM_OPT_STORAGE, _lib.PAM_STORE_REVOLVE_F16)
    eng.set_option(_lib.PAM_OPT_SNAPSHOTS, slots)
    g = eng.forward_adjoint(res).cpu().numpy()
    N = pb["g"].N
    vp = np.pad(pb["vp"].astype(np.float64), N, mode="edge")
    rn = res.cpu().numpy()
    ref = np.zeros_like(vp)
    for s in range(2):
        st = _setup(pb, s)
        states, prevs = EA.revolve_states(st, vp, slots, "f16")
        ref += EA.adjoint(st, vp, np.asarray(rn[s], dtype=np.float64), states=states, prev_states=prevs)
    ref = EA.expand_T(ref, N)
    exact, _ = _oracle_grad(pb, rn)
    e, e_exact, e_ref = rel_l2(g, ref), rel_l2(g, exact), rel_l2(ref, exact)
    print(f"  vs the restatement {e:.2e}; vs the exact float64 gradient {e_exact:.2e} "
          f"(restatement vs exact {e_ref:.2e})", flush=True)
    assert np.isfinite(g).all() and np.abs(g).max() > 0
    assert e_ref / 3 <= e_exact <= 3 * e_ref
    assert e <= 2 * e_ref


def test_batched_equals_one_shot_at_a_time():
    pb, true = _problem(0, 41, nshots=3)
    eng = make_engine(true, max_shots=3)
    dobs = eng.forward(0, 3)
    eng.set_model(pb["vp"])
    res = eng.forward(0, 3) - dobs
    g3 = eng.forward_adjoint(res).cpu().numpy()
    eng.set_option(_lib.PAM_OPT_BATCH, 1)
    g1 = eng.forward_adjoint(res).cpu().numpy()
    assert rel_l2(g1, g3) <= 1e-6


@pytest.mark.parametrize("abc", [0, 1])
def test_gradient_matches_central_differences_of_the_gpu_misfit(abc):
    """J(vp) = 1/2 |F(vp) - dobs|^2 with the GPU forward; <grad, dvp> vs
    (J(vp + h dvp) - J(vp - h dvp)) / 2h on random smooth perturbations."""
    print(f"case fd abc={abc}", flush=True)
    pb, true = _problem(abc, 50 + abc)
    eng = make_engine(true, max_shots=2)
    dobs = eng.forward(0, 2)
    vp = pb["vp"]

    def J(m):
        eng.set_model(m)
        d = (eng.forward(0, 2) - dobs).double()
        return 0.5 * float(torch.sum(d * d))

    eng.set_model(vp)
    res = eng.forward(0, 2) - dobs
    g = eng.forward_adjoint(res).cpu().numpy().astype(np.float64)
    errs = []
    for i in range(4):
        dm = smooth_field(np.random.default_rng(200 + i), *vp.shape, -1.0, 1.0).astype(np.float64)
        h = 5.0  # m/s
        fd = (J((vp + h * dm).astype(np.float32)) - J((vp - h * dm).astype(np.float32))) / (2 * h)
        an = float(np.sum(g * dm))
        errs.append(abs(fd - an) / abs(fd))
    print(f"  relative errors {['%.1e' % e for e in errs]}", flush=True)
    assert max(errs) <= 2e-3


def test_autograd_function_backpropagates_through_the_exact_adjoint():
    from pseudoacousticmodeling_amd.autograd import misfit, seismograms

    pb, true = _problem(1, 60)
    eng = make_engine(true, max_shots=2)
    dobs = eng.forward(0, 2).clone()
    vp = torch.tensor(pb["vp"], device="cuda", requires_grad=True)
    J = misfit(vp, eng, dobs)
    J.backward()
    eng.set_model(pb["vp"])
    res = eng.forward(0, 2) - dobs
    direct = eng.forward_adjoint(res)
    assert torch.equal(vp.grad, direct)
    # a loss that is linear in the gathers: grad_seis = w
    w = torch.randn(dobs.shape, device="cuda", generator=torch.Generator(device="cuda").manual_seed(3))
    vp.grad = None
    (seismograms(vp, eng) * w).sum().backward()
    np.testing.assert_allclose(vp.grad.cpu().numpy(), eng.forward_adjoint(w).cpu().numpy(), rtol=1e-6, atol=0)
    # a second forward with another model does not disturb the first one's backward
    vp2 = torch.tensor(pb["vp"] * 1.01, device="cuda", requires_grad=True)
    vp.grad = None
    J1 = misfit(vp, eng, dobs)
    _ = misfit(vp2, eng, dobs)
    J1.backward()
    assert torch.equal(vp.grad, direct)


def test_non_acoustic_physics_is_refused():
    pb = make_problem(1, 0, nz=24, nx=30, N=8, nt=60, itlag=10)
    eng = make_engine(pb)
    d = torch.zeros((1, eng.nt_data, len(pb["rz"])), device="cuda")
    with pytest.raises(_lib.PamError, match="acoustic"):
        eng.forward_adjoint(d)
