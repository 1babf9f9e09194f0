"""Source-field storage strategies of the adjoint passes (RTM image, FWI
gradient and partials): uniform checkpoints in HBM must reproduce the
save-every-step result bit for bit; float16 / bfloat16 snapshots (scaled per
snapshot by their max over the interior; the gradient stores the second time
difference formed in float32, scaled by the previous step's max) must match
the oracle's restatement of the same storage (oracle/oracle.py _SaveAll)
within 1e-4, and their distance to the float32 result -- the storage's own
error -- stays within the stated tolerance."""
import numpy as np
import pytest

from oracle.oracle import Oracle, rel_l2
from pseudoacousticmodeling_amd import _lib
from tests.problems import make_problem, make_engine
from tests.test_gpu_adjoint import _data, _oracle_adjoint

pytestmark = pytest.mark.gpu

# float16 snapshots: 11-bit significand, rounding 2^-12 relative per sample
TOL_F16 = 2e-3
# bfloat16 snapshots: 8-bit significand, rounding 2^-9 relative per sample
TOL_BF16 = 1.5e-2
# against the oracle's restatement of the same 16-bit storage: images within
# 1e-4.  A gradient reads the second time difference D_t, whose cancellation
# turns the ~1e-6 float32 differences between GPU and oracle into flipped
# 16-bit roundings at a few percent of the samples, so two correct
# implementations of the same storage differ by a fraction of that storage's
# own error: their own errors (16-bit vs float32 storage) agree within 30%,
# and they are closer to each other than 3/4 of that error.
TOL_SAME = 1e-4


def _check_same_strategy(grad, es, own_gpu, own_oracle):
    if not grad:
        assert es <= TOL_SAME
    else:
        assert abs(own_gpu / own_oracle - 1.0) <= 0.3
        assert es <= 0.75 * own_oracle


def _run(eng, data, grad, multi=False):
    if grad:
        out = eng.gradient(data, multiparameter=multi)
        return {k: v.cpu().numpy() for k, v in out.items() if v is not None}
    img, il = eng.rtm(data)
    return {"image": img.cpu().numpy(), "ilum": il.cpu().numpy()}


@pytest.mark.parametrize("phys,abc", [(p, a) for p in (0, 1, 2) for a in (0, 1)])
@pytest.mark.parametrize("grad", [False, True])
def test_checkpoints_are_bit_identical_to_save_all(phys, abc, grad):
    print(f"case checkpoint phys={phys} abc={abc} grad={grad}", flush=True)
    pb = make_problem(phys, abc, nshots=2, seed=120 + 2 * phys + abc, nt=240)
    data = _data(pb, 2, seed=12)
    eng = make_engine(pb, max_shots=2)
    multi = grad and phys >= 1
    ref = _run(eng, data, grad, multi)
    eng.set_option(_lib.PAM_OPT_STORAGE, _lib.PAM_STORE_CHECKPOINT)
    for step in (17, 0):  # 14 segments; and the default sqrt(nt)
        eng.set_option(_lib.PAM_OPT_CKPT_STEP, step)
        got = _run(eng, data, grad, multi)
        for k in ref:
            assert np.array_equal(got[k], ref[k]), f"{k} differs (checkpoint step {step})"


@pytest.mark.parametrize("phys,abc", [(0, 0), (0, 1), (1, 1), (2, 1)])
@pytest.mark.parametrize("grad", [False, True])
def test_half_precision_snapshots(phys, abc, grad):
    print(f"case half phys={phys} abc={abc} grad={grad}", flush=True)
    pb = make_problem(phys, abc, nshots=2, seed=130 + 2 * phys + abc, nt=240)
    data = _data(pb, 2, seed=13)
    eng = make_engine(pb, max_shots=2)
    key = "g_m" if grad else "image"
    ref32 = _run(eng, data, grad)[key]
    o64 = _oracle_adjoint(pb, data, fwi=grad, kind="f64")["image"]
    o32 = _oracle_adjoint(pb, data, fwi=grad, kind="f32")["image"]
    for kind, tol, name in ((_lib.PAM_STORE_ALL_F16, TOL_F16, "f16"), (_lib.PAM_STORE_ALL_BF16, TOL_BF16, "bf16")):
        eng.set_option(_lib.PAM_OPT_STORAGE, kind)
        got = _run(eng, data, grad)[key]
        e32 = rel_l2(got, ref32)
        e64 = rel_l2(got, o64)
        same = _oracle_adjoint(pb, data, fwi=grad, kind="f32", snap_kind=name)["image"]
        es = rel_l2(got, same)
        own = rel_l2(same, o32)  # the restated storage's own error
        print(f"  storage {name}: rel_l2 vs the same-strategy float32 oracle {es:.3e}; vs float32 storage "
              f"{e32:.3e} (oracle: {own:.3e}), vs FP64 oracle {e64:.3e} (float32 storage vs FP64 oracle "
              f"{rel_l2(ref32, o64):.3e})", flush=True)
        assert np.isfinite(got).all()
        _check_same_strategy(grad, es, e32, own)
        assert e32 <= tol


def test_half_precision_rejected_for_the_linearised_adjoint():
    pb = make_problem(1, 0, nshots=1, seed=140, nt=120)
    data = _data(pb, 1, seed=14)
    eng = make_engine(pb)
    eng.set_option(_lib.PAM_OPT_STORAGE, _lib.PAM_STORE_ALL_F16)
    with pytest.raises(_lib.PamError, match="PAM_E_UNSUPPORTED"):
        eng.rtm(data)


def test_half_precision_rejected_with_partials():
    """The eps / delta partials read float32 snapshots of S_t: 16-bit save-all
    snapshots with partials are refused before any launch."""
    pb = make_problem(1, 1, nshots=1, seed=141, nt=120)
    data = _data(pb, 1, seed=15)
    eng = make_engine(pb)
    for kind in (_lib.PAM_STORE_ALL_F16, _lib.PAM_STORE_ALL_BF16):
        eng.set_option(_lib.PAM_OPT_STORAGE, kind)
        with pytest.raises(_lib.PamError, match="PAM_E_UNSUPPORTED"):
            eng.gradient(data, multiparameter=True)
        eng.gradient(data)  # the plain gradient still runs
