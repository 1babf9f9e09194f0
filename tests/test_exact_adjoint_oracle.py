"""Pins of the oracle's exact discrete adjoint (oracle/exact_adjoint.py), all
in float64 on the CPU:

* its forward map equals the C oracle's acoustic forward (pam_oracle.c) and
  its CPML profiles equal the C oracle's;
* its tangent (Born) operator equals the central finite difference of the
  forward map;
* dot-product test: <T dvp, r> = <dvp, T^T r> to rounding;
* the model gradient of J = 1/2 |seis - dobs|^2 equals central finite
  differences of J along random smooth perturbations, and the interior
  gradient is the transpose of the edge-replicating expansion."""
import numpy as np
import pytest

from oracle import exact_adjoint as EA
from oracle.oracle import Oracle
from tests.problems import make_problem, oracle_forward, smooth_field


def _setup(pb, shot=0):
    g = pb["g"]
    orc = Oracle("f64")
    return EA.Setup(nzp=g.nzp, nxp=g.nxp, N=g.N, dz=g.dz, dx=g.dx, dt=g.dt, nt=g.nt, itlag=g.itlag,
                    abc=pb["abc"], src=np.asarray(pb["src"], dtype=np.float64), sz=int(pb["sz"][shot]),
                    sx=int(pb["sx"][shot]), rz=np.asarray(pb["rz"]), rx=np.asarray(pb["rx"]),
                    A=np.asarray(orc.cerjan(g.N, 0.0), dtype=np.float64), d0=pb["d0"], fpico=pb["fp"])


def _small(abc, seed, nt=140):
    return make_problem(0, abc, nz=22, nx=28, N=8, nt=nt, itlag=20, seed=seed)


@pytest.mark.parametrize("abc", [0, 1])
def test_numpy_forward_equals_the_c_oracle(abc):
    pb = _small(abc, 3)
    st = _setup(pb)
    orc = Oracle("f64")
    vp_pad = orc.expand(pb["vp"], pb["g"].N)
    ref = oracle_forward(orc, pb)["seis"]
    got = EA.forward(st, vp_pad)["seis"]
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))
    if abc == 1:
        g = pb["g"]
        ax, bx, az, bz = orc.cpml_profiles(vp_pad, g.N, g.dx, g.dz, g.dt, pb["d0"], pb["fp"])
        pr = st.profiles(np.asarray(vp_pad, dtype=np.float64))
        for a, b in ((ax, pr["ax"]), (bx, pr["bx"]), (az, pr["az"]), (bz, pr["bz"])):
            np.testing.assert_allclose(a, b, rtol=1e-9, atol=1e-15)


def test_profile_derivatives_are_the_finite_differences():
    """The CPML profiles depend on vp (Cerjan factors do not)."""
    pb = _small(1, 4)
    st = _setup(pb)
    vp = np.asarray(Oracle("f64").expand(pb["vp"], 8), dtype=np.float64)
    h = 1e-3
    p, m = st.profiles(vp + h), st.profiles(vp - h)
    pr = st.profiles(vp)
    for key in ("ax", "bx", "az", "bz"):
        fd = (p[key] - m[key]) / (2 * h)
        np.testing.assert_allclose(pr["d" + key], fd, rtol=1e-6, atol=1e-14)


@pytest.mark.parametrize("abc", [0, 1])
def test_tangent_is_the_directional_derivative(abc):
    pb = _small(abc, 5, nt=100)
    st = _setup(pb)
    vp = np.asarray(Oracle("f64").expand(pb["vp"], 8), dtype=np.float64)
    rng = np.random.default_rng(6)
    dvp = rng.standard_normal(vp.shape) * 20.0
    T = EA.tangent(st, vp, dvp)
    h = 1e-3
    fd = (EA.forward(st, vp + h * dvp)["seis"] - EA.forward(st, vp - h * dvp)["seis"]) / (2 * h)
    assert np.linalg.norm(T - fd) <= 1e-6 * np.linalg.norm(fd)


@pytest.mark.parametrize("abc", [0, 1])
def test_dot_product_adjoint(abc):
    pb = _small(abc, 7, nt=120)
    st = _setup(pb)
    vp = np.asarray(Oracle("f64").expand(pb["vp"], 8), dtype=np.float64)
    rng = np.random.default_rng(8)
    dvp = rng.standard_normal(vp.shape)
    r = rng.standard_normal((st.nt - st.itlag, len(st.rz)))
    lhs = float(np.sum(EA.tangent(st, vp, dvp) * r))
    rhs = float(np.sum(dvp * EA.adjoint(st, vp, r)))
    assert abs(lhs - rhs) <= 1e-11 * max(abs(lhs), abs(rhs))


def test_dot_product_with_duplicate_receivers_and_late_itlag():
    pb = _small(0, 9, nt=90)
    pb["rz"] = np.array([12, 12, 12, 13], dtype=np.int32)
    pb["rx"] = np.array([15, 15, 16, 20], dtype=np.int32)
    st = _setup(pb)
    st.itlag = 60
    vp = np.asarray(Oracle("f64").expand(pb["vp"], 8), dtype=np.float64)
    rng = np.random.default_rng(10)
    dvp = rng.standard_normal(vp.shape)
    r = rng.standard_normal((st.nt - st.itlag, 4))
    lhs = float(np.sum(EA.tangent(st, vp, dvp) * r))
    rhs = float(np.sum(dvp * EA.adjoint(st, vp, r)))
    assert abs(lhs - rhs) <= 1e-11 * max(abs(lhs), abs(rhs))


@pytest.mark.parametrize("abc", [0, 1])
def test_misfit_gradient_matches_central_differences(abc):
    """J(vp) = 1/2 |seis(expand(vp)) - dobs|^2 over the interior model;
    grad = expand^T adjoint(seis - dobs); checked on random smooth dvp."""
    pb = _small(abc, 11, nt=140)
    st = _setup(pb)
    N = pb["g"].N
    vp = pb["vp"].astype(np.float64)
    rng = np.random.default_rng(12)
    true = vp * (1.0 + 0.03 * rng.standard_normal(vp.shape))
    dobs = EA.forward(st, np.pad(true, N, mode="edge"))["seis"]

    def J(m):
        d = EA.forward(st, np.pad(m, N, mode="edge"))["seis"] - dobs
        return 0.5 * float(np.sum(d * d))

    res = EA.forward(st, np.pad(vp, N, mode="edge"))["seis"] - dobs
    g = EA.expand_T(EA.adjoint(st, np.pad(vp, N, mode="edge"), res), N)
    for i in range(3):
        dm = smooth_field(np.random.default_rng(100 + i), *vp.shape, -1.0, 1.0).astype(np.float64) * 10.0
        h = 1e-2
        fd = (J(vp + h * dm) - J(vp - h * dm)) / (2 * h)
        an = float(np.sum(g * dm))
        assert abs(fd - an) <= 1e-6 * abs(fd), (fd, an)


def test_expand_transpose():
    rng = np.random.default_rng(13)
    m = rng.standard_normal((5, 7))
    gp = rng.standard_normal((5 + 6, 7 + 6))
    lhs = float(np.sum(np.pad(m, 3, mode="edge") * gp))
    rhs = float(np.sum(m * EA.expand_T(gp, 3)))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
