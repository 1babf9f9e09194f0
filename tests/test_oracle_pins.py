"""Pins the CPU oracle (oracle/pam_oracle.c) against what the mathematics fixes.

The reference ships no tests, fixtures or golden vectors, and neither of its
implementations can run here (no cupy / numba / nvcc), so every oracle
function is pinned by a property that a plausible mistake (a wrong
coefficient, sign, index or transposed operand) would break:

* stencil coefficients: exactness on polynomials (which determines the 5
  second-derivative and 4 first-derivative coefficients uniquely);
* the mixed derivative and the TTI rotation: polynomial exactness, and
  rotational covariance on plane waves (TTI(theta) on k == VTI on R(theta) k);
* Sd: elliptic exactness (eps = delta -> Sd = 0), Thomsen's weak-anisotropy
  phase velocity V^2/v^2 = 1 + 2 eps n_x^4 + 2 delta n_x^2 n_z^2 + O(eps^2), and
  the exact acoustic qP phase velocity to O(t^3) as (eps, delta) = t (eps0,
  delta0), which fixes the denominator's first-order terms (a planted
  2 (1 + eps) for 2 (1 + delta) converges only at O(t^2));
* the leapfrog update: discrete energy conservation with no absorbing layer;
* physics reductions: VTI(eps=delta=0) == acoustic, TTI(theta=0) == VTI, the
  same with CPML, mirror symmetry;
* the linearised VTI / TTI adjoints: exact transposes of the forward update's
  Jacobian, the Jacobian taken by central differences of the forward oracle;
* the multiparameter partials: minus the derivative of the forward update
  with respect to eps, delta, theta, by central differences;
* the absorbing layers: energy left in the grid after the wave has left.
"""
import ctypes

import numpy as np
import pytest

from oracle.oracle import Grid, Oracle, dampening_const, rel_l2
from tests.problems import make_problem, oracle_forward

_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


F64 = Oracle("f64")


def _poly(nz, nx, h, deg, seed):
    """Random bivariate polynomial of degree <= deg in each variable on a grid
    centred at the origin, with its exact first, second and mixed derivatives."""
    rng = np.random.default_rng(seed)
    z = (np.arange(nz) - (nz - 1) / 2) * h
    x = (np.arange(nx) - (nx - 1) / 2) * h
    Z, X = np.meshgrid(z, x, indexing="ij")
    U = np.zeros_like(Z)
    Ux = np.zeros_like(Z)
    Uz = np.zeros_like(Z)
    Uxx = np.zeros_like(Z)
    Uzz = np.zeros_like(Z)
    Uxz = np.zeros_like(Z)
    from math import factorial

    for a in range(deg + 1):
        for b in range(deg + 1):
            c = rng.standard_normal() / (factorial(a) * factorial(b))
            U += c * X ** a * Z ** b
            if a >= 1:
                Ux += c * a * X ** (a - 1) * Z ** b
            if b >= 1:
                Uz += c * b * X ** a * Z ** (b - 1)
            if a >= 2:
                Uxx += c * a * (a - 1) * X ** (a - 2) * Z ** b
            if b >= 2:
                Uzz += c * b * (b - 1) * X ** a * Z ** (b - 2)
            if a >= 1 and b >= 1:
                Uxz += c * a * b * X ** (a - 1) * Z ** (b - 1)
    return dict(U=U, Ux=Ux, Uz=Uz, Uxx=Uxx, Uzz=Uzz, Uxz=Uxz)


IN = (slice(4, -4), slice(4, -4))


def test_second_derivative_exact_on_degree9_polynomials():
    n, h = 24, 0.3
    f = _poly(n, n, h, 9, seed=1)
    Uf = np.zeros((n, n))
    vp = np.ones((n, n))
    F64.lib.or_update_acoustic(_p(Uf), _p(f["U"]), _p(vp), _I(n), _I(n), _D(h), _D(h), _D(1.0))
    lap = (Uf - 2 * f["U"])[IN]
    exact = (f["Uxx"] + f["Uzz"])[IN]
    assert rel_l2(lap, exact) < 1e-9
    # a stencil of lower order is not exact: the test has teeth
    g = _poly(n, n, h, 11, seed=2)
    Uf[...] = 0
    F64.lib.or_update_acoustic(_p(Uf), _p(g["U"]), _p(vp), _I(n), _I(n), _D(h), _D(h), _D(1.0))
    assert rel_l2((Uf - 2 * g["U"])[IN], (g["Uxx"] + g["Uzz"])[IN]) > 1e-9


def test_first_derivative_exact_on_degree8_polynomials():
    # psi = a psi + b dU with a = 0, b = 1 gives dU/dx (x strips) and dU/dz (z
    # strips); N = n/2 makes the strips cover the whole stencil interior
    n, h = 24, 0.3
    N = n // 2
    f = _poly(n, n, h, 8, seed=3)
    zero = np.zeros((n, n))
    one = np.ones((n, n))
    psix = np.zeros((n, n))
    psiz = np.zeros((n, n))
    F64.lib.or_cpml_psi(_p(psix), _p(psiz), _p(f["U"]), _p(zero), _p(one), _p(zero), _p(one), _I(n), _I(n),
                        _I(N), _D(h), _D(h))
    assert rel_l2(psix[IN], f["Ux"][IN]) < 1e-10
    assert rel_l2(psiz[IN], f["Uz"][IN]) < 1e-10


def test_mixed_derivative_exact_and_tti_coefficients():
    # eps = delta (so Sd = 0), theta = 45 deg: P = (1+e) (Uxx + Uzz) - 2 e Uxz
    n, h, e = 24, 0.3, 0.2
    f = _poly(n, n, h, 8, seed=4)
    Uf = np.zeros((n, n))
    one = np.ones((n, n))
    eps = np.full((n, n), e)
    th = np.full((n, n), np.pi / 4)
    F64.lib.or_update_tti(_p(Uf), _p(f["U"]), _p(one), _p(eps), _p(eps), _p(th), _I(n), _I(n), _D(h), _D(h),
                          _D(1.0))
    P = (Uf - 2 * f["U"])[IN]
    exact = ((1 + e) * (f["Uxx"] + f["Uzz"]) - 2 * e * f["Uxz"])[IN]
    assert rel_l2(P, exact) < 1e-9


def _plane(n, h, kx, kz, phase=0.3):
    z = np.arange(n) * h
    x = np.arange(n) * h
    Z, X = np.meshgrid(z, x, indexing="ij")
    ph = kx * X + kz * Z + phase
    return np.cos(ph), np.sin(ph)


def _ratio(phys, kx, kz, eps, dl, th=0.0, n=40, h=10.0, lib=None):
    """-P/(k^2 U) for a plane wave: the normalised squared phase velocity."""
    lib = F64.lib if lib is None else lib
    U, S = _plane(n, h, kx, kz)
    Uf = np.zeros((n, n))
    one = np.ones((n, n))
    E = np.full((n, n), eps)
    Dl = np.full((n, n), dl)
    if phys == 1:
        lib.or_update_vti(_p(Uf), _p(U), _p(one), _p(E), _p(Dl), _I(n), _I(n), _D(h), _D(h), _D(1.0))
    else:
        T = np.full((n, n), th)
        lib.or_update_tti(_p(Uf), _p(U), _p(one), _p(E), _p(Dl), _p(T), _I(n), _I(n), _D(h), _D(h), _D(1.0))
    P = Uf - 2 * U
    m = np.zeros_like(U, dtype=bool)
    m[IN] = True
    m &= (np.abs(U) > 0.4) & (np.abs(S) > 0.4)
    return (-P[m] / ((kx * kx + kz * kz) * U[m]))


def test_sd_elliptic_and_weak_anisotropy_limits():
    k = 2 * np.pi / 600.0  # 60 cells per wavelength at h = 10
    for ang in (10.0, 35.0, 60.0, 80.0):
        nx_, nz_ = np.sin(np.deg2rad(ang)), np.cos(np.deg2rad(ang))
        # elliptic: Sd = 0 and V^2/v^2 = (1+2e) nx^2 + nz^2
        R = _ratio(1, k * nx_, k * nz_, 0.2, 0.2)
        assert np.allclose(R, (1 + 0.4) * nx_ ** 2 + nz_ ** 2, rtol=1e-7, atol=0)
        # weak anisotropy: V^2/v^2 - 1 = 2 e nx^4 + 2 d nx^2 nz^2 + O(e^2)
        e, d = 2e-4, 1e-4
        R = _ratio(1, k * nx_, k * nz_, e, d)
        first = 2 * e * nx_ ** 4 + 2 * d * nx_ ** 2 * nz_ ** 2
        assert np.all(np.abs(R - 1 - first) < 20 * e * e + 1e-9)
        # a wrong sign of the Sd numerator would be first order
        assert abs(2 * (e - d) * nx_ ** 2 * nz_ ** 2) > 20 * e * e


def _qp_exact(e, d, nx_, nz_):
    """Squared qP phase velocity over vp^2 of the acoustic (shear-wave-free)
    VTI medium: with A = 1 + 2 e nx^2 and Y = 2 (e - d) nx^2 nz^2 for a unit
    wave normal, V^2/vp^2 = A/2 + sqrt(A^2 - 4 Y)/2 (Alkhalifah's acoustic
    approximation, the Vs -> 0 limit of Thomsen's exact qP velocity)."""
    A = 1.0 + 2.0 * e * nx_ ** 2
    Y = 2.0 * (e - d) * nx_ ** 2 * nz_ ** 2
    return 0.5 * A + 0.5 * np.sqrt(A * A - 4.0 * Y)


def _sd_orders(lib):
    """log2 of the error ratio between (e, d) = t (0.2, 0.05) and t/2 against the
    exact qP velocity, per angle: about 3 when the pseudo-acoustic velocity
    A - Y / (A - Y) (the Sd operator on a plane wave) agrees with it to
    second order, about 2 for a denominator wrong at first order."""
    k = 2 * np.pi / 1200.0  # 120 cells per wavelength: the stencil's own error is ~ (k h)^8 ~ 5e-11
    e0, d0 = 0.2, 0.05
    orders = []
    for ang in (20.0, 45.0, 70.0):
        nx_, nz_ = np.sin(np.deg2rad(ang)), np.cos(np.deg2rad(ang))
        errs = []
        for t in (0.2, 0.1, 0.05):
            R = np.median(_ratio(1, k * nx_, k * nz_, t * e0, t * d0, n=48, lib=lib))
            errs.append(abs(R - _qp_exact(t * e0, t * d0, nx_, nz_)))
        orders.append([np.log2(errs[i] / errs[i + 1]) for i in range(2)])
        # and the error is small in absolute terms: O(t^3) with its known coefficient scale
        assert errs[-1] < 1e-3
    return np.array(orders)


def test_sd_matches_the_exact_qp_velocity_to_second_order(tmp_path):
    """The Sd denominator beyond first order in (eps, delta): the oracle's plane-
    wave velocity converges to the exact acoustic qP velocity at O(t^3) as
    (eps, delta) = t (eps0, delta0) -> 0; an oracle built with 2 (1 + eps) in
    place of 2 (1 + delta) in the denominator converges only at O(t^2)."""
    good = _sd_orders(F64.lib)
    assert np.all(good > 2.8), good
    import os
    import subprocess

    src = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "pam_oracle.c")
    so = str(tmp_path / "liboracle_planted.so")
    subprocess.run(["gcc", "-O2", "-fPIC", "-shared", "-std=c11", "-ffp-contract=off", "-DORACLE_F64",
                    "-DORACLE_PLANT_SD_DEN_EPS", "-o", so, src, "-lm"], check=True)
    bad = _sd_orders(ctypes.CDLL(so))
    assert np.all(bad < 2.2), bad


def test_tti_is_vti_in_the_rotated_frame():
    k = 2 * np.pi / 600.0
    e, d = 0.2, 0.05
    for th_deg, ang in ((25.0, 10.0), (-35.0, 50.0), (60.0, -20.0)):
        th = np.deg2rad(th_deg)
        kx, kz = k * np.sin(np.deg2rad(ang)), k * np.cos(np.deg2rad(ang))
        kxi = np.cos(th) * kx - np.sin(th) * kz
        keta = np.sin(th) * kx + np.cos(th) * kz
        Rt = _ratio(2, kx, kz, e, d, th)
        Rv = _ratio(1, kxi, keta, e, d)
        assert abs(np.median(Rt) - np.median(Rv)) < 1e-7
        assert np.ptp(Rt) < 1e-7
        # the unrotated VTI value differs: the rotation matters
        assert abs(np.median(_ratio(1, kx, kz, e, d)) - np.median(Rt)) > 1e-3


def test_leapfrog_energy_conserved_without_absorbing_layer():
    g = Grid(nz=40, nx=44, N=0, dz=10.0, dx=10.0, dt=0.001, nt=0)
    rng = np.random.default_rng(5)
    vp = (1500 + 1000 * rng.random((g.nzp, g.nxp))).astype(np.float64)
    A = np.zeros(1)
    cur = np.zeros((g.nzp, g.nxp))
    prev = np.zeros_like(cur)
    cur[20, 22] = 1.0
    one = np.ones_like(vp)

    def L(u):
        out = np.zeros_like(u)
        F64.lib.or_update_acoustic(_p(out), _p(u), _p(one), _I(g.nzp), _I(g.nxp), _D(g.dz), _D(g.dx), _D(1.0))
        return out - 2 * u * (np.pad(np.ones((g.nzp - 8, g.nxp - 8)), 4) > 0)

    M = 1.0 / vp ** 2
    energies = []
    for k in range(300):
        F64.forward_step(0, 0, prev, cur, g, (vp, None, None, None), A, None, -1, -1, 0.0)
        nxt = prev  # updated in place
        D = (nxt - cur) / g.dt
        energies.append(np.sum(M * D * D) - np.sum(nxt * L(cur)))
        prev, cur = cur, nxt
    energies = np.array(energies)
    assert np.max(np.abs(energies - energies[0])) < 1e-9 * abs(energies[0])


@pytest.mark.parametrize("abc", [0, 1])
def test_vti_without_anisotropy_is_acoustic(abc):
    pb = make_problem(1, abc, seed=21, nt=200)
    pa = dict(pb, phys=0)
    pz = dict(pb, eps=np.zeros_like(pb["eps"]), dl=np.zeros_like(pb["dl"]))
    a = oracle_forward(F64, pa)["seis"]
    v = oracle_forward(F64, pz)["seis"]
    assert rel_l2(v, a) < 1e-12
    assert rel_l2(oracle_forward(F64, pb)["seis"], a) > 1e-3


@pytest.mark.parametrize("abc", [0, 1])
def test_tti_with_zero_tilt_is_vti(abc):
    pb = make_problem(2, abc, seed=22, nt=200)
    p0 = dict(pb, th=np.zeros_like(pb["th"]))
    pv = dict(pb, phys=1)
    assert rel_l2(oracle_forward(F64, p0)["seis"], oracle_forward(F64, pv)["seis"]) < 1e-12
    assert rel_l2(oracle_forward(F64, pb)["seis"], oracle_forward(F64, pv)["seis"]) > 1e-3


@pytest.mark.parametrize("phys,abc", [(0, 0), (1, 1), (2, 0)])
def test_mirror_symmetry(phys, abc):
    pb = make_problem(phys, abc, nx=57, seed=23, nt=200)
    for key in ("vp", "eps", "dl"):
        if pb[key] is not None:
            col = pb[key][:, :1]
            pb[key] = np.repeat(col, pb["vp"].shape[1], axis=1).astype(np.float32)
    if phys == 2:
        pb["th"] = np.zeros_like(pb["th"])
    g = pb["g"]
    pb["sx"] = np.array([g.N + g.nx // 2], dtype=np.int32)
    pb["rx"] = np.arange(g.N, g.N + g.nx, dtype=np.int32)
    pb["rz"] = np.full_like(pb["rx"], g.N + 2)
    s = oracle_forward(F64, pb)["seis"]
    assert rel_l2(s[:, ::-1], s) < 1e-12


def _fd_jacobian_dot(phys, S, x, y, model, h):
    """<J x, y> with J the Jacobian of the forward update at S (central differences)."""
    n = S.shape[0]
    vp, eps, dl, th = model
    a = 1e-6

    def F(U):
        out = np.zeros_like(U)
        if phys == 1:
            F64.lib.or_update_vti(_p(out), _p(U), _p(vp), _p(eps), _p(dl), _I(n), _I(n), _D(h), _D(h), _D(1.0))
        else:
            F64.lib.or_update_tti(_p(out), _p(U), _p(vp), _p(eps), _p(dl), _p(th), _I(n), _I(n), _D(h), _D(h),
                                  _D(1.0))
        return out

    Jx = (F(S + a * x) - F(S - a * x)) / (2 * a)
    return float(np.sum(Jx * y))


@pytest.mark.parametrize("phys", [1, 2])
def test_linearised_adjoint_is_the_transpose_of_the_forward_jacobian(phys):
    n, h = 30, 1.0
    rng = np.random.default_rng(30 + phys)
    inner = np.zeros((n, n))
    inner[IN] = 1.0
    S = rng.standard_normal((n, n))
    x = rng.standard_normal((n, n)) * inner
    y = rng.standard_normal((n, n)) * inner
    vp = np.full((n, n), 0.3)  # constant: the reference keeps v^2 outside the derivatives
    eps = 0.25 * rng.random((n, n))
    dl = np.minimum(0.1 * rng.random((n, n)), eps)
    th = rng.uniform(-0.6, 0.6, (n, n))
    lhs = _fd_jacobian_dot(phys, S, x, y, (vp, eps, dl, th), h)
    adj = np.zeros((n, n))
    if phys == 1:
        F64.lib.or_adjoint_vti(_p(adj), _p(y), _p(S), _p(vp), _p(eps), _p(dl), _I(n), _I(n), _D(h), _D(h), _D(1.0))
    else:
        F64.lib.or_adjoint_tti(_p(adj), _p(y), _p(S), _p(vp), _p(eps), _p(dl), _p(th), _I(n), _I(n), _D(h), _D(h),
                               _D(1.0))
    rhs = float(np.sum(x * adj))
    assert abs(lhs - rhs) < 1e-7 * abs(lhs)


@pytest.mark.parametrize("phys", [1, 2])
def test_multiparameter_partials_are_minus_the_update_derivative(phys):
    n, h = 30, 2.0
    rng = np.random.default_rng(40 + phys)
    S = rng.standard_normal((n, n))
    one = np.ones((n, n))
    eps = 0.25 * rng.random((n, n))
    dl = np.minimum(0.1 * rng.random((n, n)), eps)
    th = rng.uniform(-0.6, 0.6, (n, n))

    def P(e, d, t):
        out = np.zeros((n, n))
        if phys == 1:
            F64.lib.or_update_vti(_p(out), _p(S), _p(one), _p(e), _p(d), _I(n), _I(n), _D(h), _D(h), _D(1.0))
        else:
            F64.lib.or_update_tti(_p(out), _p(S), _p(one), _p(e), _p(d), _p(t), _I(n), _I(n), _D(h), _D(h),
                                  _D(1.0))
        return out - 2 * S

    a = 1e-6
    fd_e = -(P(eps + a, dl, th) - P(eps - a, dl, th)) / (2 * a)
    fd_d = -(P(eps, dl + a, th) - P(eps, dl - a, th)) / (2 * a)
    fd_t = -(P(eps, dl, th + a) - P(eps, dl, th - a)) / (2 * a)
    adj = np.ones((n, n))
    ge = np.zeros((n, n))
    gd = np.zeros((n, n))
    gt = np.zeros((n, n))
    if phys == 1:
        F64.lib.or_grad_vti(_p(S), _p(adj), _p(ge), _p(gd), _I(n), _I(n), _D(h), _D(h), _p(eps), _p(dl))
    else:
        F64.lib.or_grad_tti(_p(S), _p(adj), _p(ge), _p(gd), _p(gt), _I(n), _I(n), _D(h), _D(h), _p(eps), _p(dl),
                            _p(th))
    assert rel_l2(ge[IN], fd_e[IN]) < 1e-6
    assert rel_l2(gd[IN], fd_d[IN]) < 1e-6
    if phys == 2:
        assert rel_l2(gt[IN], fd_t[IN]) < 1e-6


@pytest.mark.parametrize("abc", [0, 1])
def test_absorbing_layer_absorbs(abc):
    # homogeneous model; after the wave has crossed the boundary, the energy
    # left in the interior is a small fraction of the peak
    g = Grid(nz=60, nx=60, N=20, dz=10.0, dx=10.0, dt=0.001, nt=900)
    vp = np.full((g.nz, g.nx), 2000.0, dtype=np.float32)
    from oracle.oracle import ricker_cpp

    src = ricker_cpp(g.nt, g.dt, 30.0, 2 * np.sqrt(np.pi) / 30.0, g.dx, g.dz)
    model = (F64.expand(vp, g.N), None, None, None)
    d0, fp = dampening_const(g.N, 30.0)

    def left(abc_, A):
        out = {}
        for nt in (250, 900):
            r = F64.forward_shot(0, abc_, g, model, src, g.N + 30, g.N + 30, [g.N], [g.N], A=A, cpml=(d0, fp),
                                 nsteps=nt)
            out[nt] = float(np.sum(r["current"][g.N:-g.N, g.N:-g.N] ** 2))
        return out[900] / out[250]

    rigid = left(0, np.ones(g.N))  # no damping: the rim reflects everything back
    assert rigid > 0.1
    # 20 cells: Cerjan keeps < 1% of the energy, CPML < 0.1%
    if abc == 0:
        assert left(0, F64.cerjan(g.N)) < 1e-2
    else:
        assert left(1, None) < 1e-3


def test_cpml_profiles_closed_form():
    g = Grid(nz=10, nx=12, N=10, dz=10.0, dx=12.5, dt=0.001, nt=1)
    vp = np.full((g.nzp, g.nxp), 2500.0)
    d0, fp = dampening_const(g.N, 30.0)
    assert abs(d0 - 3 * np.log(1e3)) < 1e-12  # R = 1e-3 for N = 10, M = 2
    ax, bx, az, bz = F64.cpml_profiles(vp, g.N, g.dx, g.dz, g.dt, d0, fp)
    # no damping in the interior
    assert np.all(ax[:, g.N:g.nxp - g.N] == 1.0) and np.all(bx[:, g.N:g.nxp - g.N] == 0.0)
    assert np.all(az[g.N:g.nzp - g.N, :] == 1.0) and np.all(bz[g.N:g.nzp - g.N, :] == 0.0)
    # outermost cell: pos = (N-1)/N, d = d0 pos^2 vp / (2 N h), alpha = pi fp (1 - pos^2)
    pos = (g.N - 1) / g.N
    d = d0 * pos * pos * 2500.0 / (2 * g.N * g.dx)
    al = np.pi * fp * (1 - pos * pos)
    a = np.exp(-(d + al) * g.dt)
    assert abs(ax[3, 0] - a) < 1e-9 and abs(bx[3, 0] - d / (d + al) * (a - 1)) < 1e-9
    # mirror symmetry of the left / right and top / bottom strips
    assert np.array_equal(ax, ax[:, ::-1]) and np.array_equal(az, az[::-1, :])
    # damping grows towards the outer edge
    assert np.all(np.diff(ax[5, :g.N]) > 0)


def test_cerjan_taper_closed_form():
    A = F64.cerjan(50)
    i = np.arange(50)
    assert np.allclose(A, np.exp(-((50 - i) / (1.4142 * 150)) ** 2), rtol=1e-14)
    B = F64.cerjan(50, 0.0045)
    assert np.allclose(B, np.exp(-(0.0045 * (50 - i)) ** 2), rtol=1e-14)


@pytest.mark.parametrize("phys,abc", [(p, a) for p in (0, 1, 2) for a in (0, 1)])
def test_float32_build_tracks_float64(phys, abc):
    pb = make_problem(phys, abc, seed=50 + 2 * phys + abc, nt=300)
    s32 = oracle_forward(Oracle("f32"), pb)["seis"]
    s64 = oracle_forward(F64, pb)["seis"]
    assert rel_l2(s32, s64) < 1e-4
