"""The per-step adjoint interface (pam_adjoint_reset / pam_adjoint_step, the
mirror of Migration::backward_step, cpp/Migration.cu:238, and
migration.backward_step[GPU], src/Migration.py:237-286): one shot driven step
by step -- the source field from the per-step forward interface, then the
adjoint steps t = nt-2 .. 1 imaging into caller buffers -- gives pam_rtm's
image bit for bit (normalisation off, the same kernel variant), for every
physics and absorbing boundary, also with the CPML memory of the forward pass
kept; the illumination is the sum of S_t^2 over the steps driven."""
import numpy as np
import pytest
import torch

from pseudoacousticmodeling_amd import _lib
from tests.problems import make_problem, make_engine
from tests.test_gpu_adjoint import _data

pytestmark = pytest.mark.gpu


def _forward_fields(eng, shot):
    """S_k for k = 0 .. nt-1 from the per-step forward interface."""
    eng.reset(shot)
    S = []
    for k in range(eng.nt):
        S.append(eng.get_field(0).clone())
        eng.forward_step(k)
    return S


@pytest.mark.parametrize("phys,abc,keep_cpml", [(p, a, False) for p in (0, 1, 2) for a in (0, 1)] +
                         [(p, 1, True) for p in (0, 1, 2)])
def test_adjoint_steps_give_pam_rtm_bit_for_bit(phys, abc, keep_cpml):
    print(f"case adjoint step phys={phys} abc={abc} keep_cpml={keep_cpml}", flush=True)
    pb = make_problem(phys, abc, nshots=2, seed=300 + 2 * phys + abc, nt=200)
    data = _data(pb, 2, seed=31)
    eng = make_engine(pb, max_shots=2)
    eng.set_option(_lib.PAM_OPT_TILE, 0)
    eng.set_option(_lib.PAM_OPT_NORMALIZE, 0)
    eng.set_option(_lib.PAM_OPT_BATCH, 1)
    if keep_cpml:
        eng.set_option(_lib.PAM_OPT_RESET_ADJ_CPML, 0)
    shot = 1
    ref_img, _ = eng.rtm(data[shot:shot + 1], shot0=shot, nshots=1)
    ref_img = ref_img.cpu().numpy()
    S = _forward_fields(eng, shot)
    eng.adjoint_reset(shot, data[shot])
    image = torch.zeros((eng.nz, eng.nx), dtype=torch.float32, device=eng.device)
    ilum = torch.zeros_like(image)
    for t in range(eng.nt - 2, 0, -1):
        eng.adjoint_step(t, S[t], image, ilum)
    img = image.cpu().numpy()
    assert np.isfinite(img).all() and np.abs(img).max() > 0
    assert np.array_equal(img, ref_img)
    N = eng.N
    ilum_ref = sum((S[t][N:N + eng.nz, N:N + eng.nx].double() ** 2) for t in range(1, eng.nt - 1)).cpu().numpy()
    assert np.allclose(ilum.cpu().numpy(), ilum_ref, rtol=1e-5, atol=1e-6 * ilum_ref.max())
    eng.close()


def test_adjoint_field_readback_and_order():
    pb = make_problem(1, 1, nshots=1, seed=320, nt=80)
    data = _data(pb, 1, seed=32)
    eng = make_engine(pb)
    with pytest.raises(_lib.PamError, match="PAM_E_STATE"):
        eng.adjoint_step(10)
    eng.adjoint_reset(0, data[0])
    lam = eng.get_field(2)  # the first step's field: data row t injected by the step itself
    assert float(lam.abs().max()) == 0.0
    eng.adjoint_step(eng.nt - 1)
    with pytest.raises(_lib.PamError, match="PAM_E_STATE"):
        eng.adjoint_step(eng.nt - 1)  # not decreasing
    eng.adjoint_step(eng.nt - 2)
    assert float(eng.get_field(2).abs().max()) > 0
    eng.rtm(data)  # a whole run reuses the second field set
    with pytest.raises(_lib.PamError, match="PAM_E_STATE"):
        eng.adjoint_step(eng.nt - 3)
    eng.close()
