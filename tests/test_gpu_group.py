"""Groups of grids (pam_group_forward): models of different sizes, each with
its own shots, advanced in one launch per time step must give, bit for bit,
the seismograms each model's own forward run gives with the same variant --
the 16-row tile (with the group's own CPML psi pass) or a z-march -- and match
the oracle."""
import numpy as np
import pytest

from oracle.oracle import Oracle, rel_l2
from pseudoacousticmodeling_amd import _lib
from tests.problems import make_problem, make_engine, oracle_forward

pytestmark = pytest.mark.gpu

SIZES = [(40, 56), (52, 40), (32, 76)]


@pytest.mark.parametrize("phys,abc", [(0, 0), (1, 1), (2, 1)])
@pytest.mark.parametrize("graph", [1, 0])
@pytest.mark.parametrize("tile", [5, 0])
def test_group_forward_equals_each_model_alone(phys, abc, graph, tile):
    from pseudoacousticmodeling_amd.engine import Group

    print(f"case group phys={phys} abc={abc} graph={graph} tile={tile}", flush=True)
    shots = [2, 1, 3]
    pbs = [make_problem(phys, abc, nz=nz, nx=nx, nshots=k, seed=700 + 10 * i + phys, nt=181)
           for i, ((nz, nx), k) in enumerate(zip(SIZES, shots))]
    engs = [make_engine(pb, max_shots=k) for pb, k in zip(pbs, shots)]
    for e in engs:
        e.set_option(_lib.PAM_OPT_TILE, tile)
        e.set_option(_lib.PAM_OPT_GRAPH, graph)
    refs = [e.forward(0, k).cpu().numpy() for e, k in zip(engs, shots)]
    grp = Group(engs)
    for call in range(2):
        outs = [o.cpu().numpy() for o in grp.forward(shots)]
        for i, (o, r) in enumerate(zip(outs, refs)):
            assert np.array_equal(o, r), f"member {i} differs (call {call})"
    assert grp.tile() == tile
    orc = oracle_forward(Oracle("f32"), pbs[2], shot=2)["seis"]
    assert rel_l2(outs[2][2], orc) <= 1e-4
    grp.close()


def test_group_autotunes_a_variant():
    from pseudoacousticmodeling_amd.engine import Group

    pbs = [make_problem(1, 1, nz=nz, nx=nx, nshots=2, seed=750 + i, nt=120) for i, (nz, nx) in enumerate(SIZES)]
    engs = [make_engine(pb, max_shots=2) for pb in pbs]
    grp = Group(engs)
    outs = grp.forward([2, 2, 2])
    assert grp.tile() in (0, 4, 5, 6)
    for pb, o in zip(pbs, outs):
        orc = oracle_forward(Oracle("f32"), pb, shot=1)["seis"]
        assert rel_l2(o[1].cpu().numpy(), orc) <= 1e-4


def test_group_rejects_mismatched_members():
    from pseudoacousticmodeling_amd.engine import Group

    a = make_engine(make_problem(0, 0, nshots=1, seed=760, nt=100))
    b = make_engine(make_problem(0, 1, nshots=1, seed=761, nt=100))
    with pytest.raises(_lib.PamError, match="PAM_E_ARG"):
        Group([a, b])


@pytest.mark.parametrize("lanes", [2, 1])
def test_group_in_sub_groups_equals_each_model_alone(lanes):
    """A small PAM_OPT_GROUP_BYTES splits the group into sub-groups (one member
    each here); with two lanes they alternate between two streams."""
    from pseudoacousticmodeling_amd.engine import Group

    print(f"case group sub-groups lanes={lanes}", flush=True)
    shots = [2, 1, 3]
    pbs = [make_problem(1, 1, nz=nz, nx=nx, nshots=k, seed=780 + i, nt=161)
           for i, ((nz, nx), k) in enumerate(zip(SIZES, shots))]
    engs = [make_engine(pb, max_shots=k) for pb, k in zip(pbs, shots)]
    for e in engs:
        e.set_option(_lib.PAM_OPT_TILE, 0)
    refs = [e.forward(0, k).cpu().numpy() for e, k in zip(engs, shots)]
    engs[0].set_option(_lib.PAM_OPT_GROUP_BYTES, 1)
    engs[0].set_option(_lib.PAM_OPT_LANES, lanes)
    assert engs[0].get_option(_lib.PAM_OPT_GROUP_BYTES) == 1
    grp = Group(engs)
    for call in range(2):
        outs = [o.cpu().numpy() for o in grp.forward(shots)]
        for i, (o, r) in enumerate(zip(outs, refs)):
            assert np.array_equal(o, r), f"member {i} differs (call {call})"
    grp.close()


@pytest.mark.parametrize("storage", [_lib.PAM_STORE_CHECKPOINT, _lib.PAM_STORE_ALL_F16, _lib.PAM_STORE_REVOLVE_F16])
def test_group_gradient_equals_each_model_alone(storage):
    """pam_group_gradient: every member's g_m is the one its own pam_gradient
    gives, bit for bit, with its storage option (float16-scaled snapshots or
    revolve slots included), on a second call too (the run scratch released
    between members is allocated again); each |g_m|^2 is the float64 sum of
    squares of that gradient."""
    import torch

    from pseudoacousticmodeling_amd.engine import Group
    from tests.test_gpu_adjoint import _data

    print(f"case group gradient storage={storage}", flush=True)
    shots = [2, 1, 3]
    pbs = [make_problem(1, 1, nz=nz, nx=nx, nshots=k, seed=790 + i, nt=181)
           for i, ((nz, nx), k) in enumerate(zip(SIZES, shots))]
    engs = [make_engine(pb, max_shots=k) for pb, k in zip(pbs, shots)]
    res = [torch.as_tensor(_data(pb, k, seed=33 + i), device="cuda")
           for i, (pb, k) in enumerate(zip(pbs, shots))]
    for e in engs:
        e.set_option(_lib.PAM_OPT_STORAGE, storage)
    refs = [e.gradient(r)["g_m"].cpu().numpy() for e, r in zip(engs, res)]
    grp = Group(engs)
    for call in range(2):
        outs, norms = grp.gradient(res)
        for i, (o, ref, nrm) in enumerate(zip(outs, refs, norms)):
            o = o.cpu().numpy()
            assert np.abs(o).max() > 0
            assert np.array_equal(o, ref), f"member {i} differs (call {call})"
            n64 = float(np.sum(o.astype(np.float64) ** 2))
            assert abs(nrm - n64) <= 1e-12 * n64, (nrm, n64)
    grp.close()
