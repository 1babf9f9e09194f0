"""GPU preprocessing of the drivers (pam_smooth_model, pam_mute) against the
oracle's restatements of the reference rules: the Python driver's 2-D
Gaussian smoothing with edge replication and the water column restored
(src/utils.py:76-129), the C++ driver's separable clamped smoothing
(cpp/Migration.cu:147-236), and the two top-mute rules (src/utils.py:30-49 in
float64, cpp/Migration.cu:86-115 in float32).  The smoothing weights come from
exp() on the host, where numpy and the C library may differ in the last bit,
so the smoothed models are compared to 1e-6 relative; the mute has no such
step and is compared bit for bit."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from pseudoacousticmodeling_amd import _lib
from pseudoacousticmodeling_amd.engine import mute, smooth_model
from tests.problems import smooth_field

pytestmark = pytest.mark.gpu


def _model(seed, nz=64, nx=96):
    rng = np.random.default_rng(seed)
    vp = smooth_field(rng, nz, nx, 1800.0, 4000.0)
    vp[:9, :] = 1500.0          # a water layer
    vp[9:14, 40:60] = 1500.0    # water reaching deeper in some columns
    return vp


@pytest.mark.parametrize("sigma", [1.5, 3.0])
@pytest.mark.parametrize("parameter", [False, True])
def test_smooth_python_rule(sigma, parameter):
    vp = _model(900)
    f = (vp / 1e4).astype(np.float32) if parameter else vp
    water = np.abs(vp - vp.min()) < 1e-3
    got = smooth_model(f, sigma, water, parameter)
    ref = O.smooth_model_py(f, sigma, water, parameter=parameter)
    err = np.max(np.abs(got - ref) / np.abs(ref))
    print(f"  python rule sigma={sigma} parameter={parameter}: max rel {err:.2e}", flush=True)
    assert err <= 1e-6
    # untouched above the first dry cell of each column
    assert np.array_equal(got[:9], f[:9])


@pytest.mark.parametrize("sigma", [1.5, 3.0])
@pytest.mark.parametrize("parameter", [False, True])
def test_smooth_cpp_rule(sigma, parameter):
    vp = _model(901)
    f = (vp / 1e4).astype(np.float32) if parameter else vp
    mask = np.abs(vp - vp.min()) < 1e-3
    got = smooth_model(f, sigma, mask, parameter, mode=_lib.PAM_PREP_CPP)
    ref = O.smooth_model_cpp(f, sigma, mask, parameter=parameter)
    err = np.max(np.abs(got - ref) / np.abs(ref))
    print(f"  C++ rule sigma={sigma} parameter={parameter}: max rel {err:.2e}", flush=True)
    assert err <= 1e-6
    assert np.array_equal(got[mask], f[mask])


def test_smooth_rejects_a_kernel_larger_than_the_model():
    vp = _model(902, nz=12, nx=40)
    with pytest.raises(_lib.PamError, match="does not fit"):
        smooth_model(vp, 3.0, None)


@pytest.mark.parametrize("rule", ["py", "cpp"])
def test_mute_rules(rule):
    rng = np.random.default_rng(903)
    nshots, nt, nrec = 3, 400, 57
    seis = rng.standard_normal((nshots, nt, nrec)).astype(np.float32)
    rec_x = np.arange(nrec) * 12.5
    rec_z = np.full(nrec, 10.0)
    sx, sz = [100.0, 350.0, 600.0], [10.0, 12.0, 10.0]
    d = torch.from_numpy(seis.copy()).cuda()
    mode = _lib.PAM_PREP_PY if rule == "py" else _lib.PAM_PREP_CPP
    mute(d, rec_x, rec_z, sx, sz, 0.001, 0.04, 0.05, 1500.0, mode=mode)
    got = d.cpu().numpy()
    for s in range(nshots):
        if rule == "py":
            ref = O.mute(seis[s], rec_x, rec_z, sx[s], sz[s], 0.001, 0.04, 0.05, 1500.0)
        else:
            ref = O.mute_cpp(seis[s], rec_x, rec_z, sx[s], sz[s], 0.001, 0.04, 0.05, 1500.0)
        assert np.array_equal(got[s], ref), f"shot {s}"
