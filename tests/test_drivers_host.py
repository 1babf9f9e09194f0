"""Host side of the drop-in drivers: Parameters.json / CSV parsing and the
derived sizes (src/survey.py:5-141, cpp/Survey.cpp:20-176), the C++ CLI's
parse (``--check``), the host preprocessing against the oracle (mute,
smoothing) and the L-BFGS two-loop recursion (src/Inversion.py:88-123)."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as orc_mod
from pseudoacousticmodeling_amd import prep
from pseudoacousticmodeling_amd.inversion import two_loop_recursion
from pseudoacousticmodeling_amd.survey import parameters, read_geometry_csv
from tests.surveys import write_csv, write_survey

BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "pseudoacousticmodeling_amd", "bin")


def test_parameters_reads_every_key_and_the_derived_sizes(tmp_path):
    path, models = write_survey(str(tmp_path), phys=1, abc=1, nshots=3)
    p = parameters(path)
    assert p.approximation == "VTI" and p.ABC == "CPML" and p.migration == "onthefly"
    assert (p.dx, p.dz, p.dt, p.L, p.D, p.T, p.N_abc, p.fcut) == (10.0, 10.0, 0.001, 550.0, 390.0, 0.3, 12, 30.0)
    # Python driver: int() truncation of tlag / dt = 118.16
    assert p.nt_data == 301 and p.itlag == 118 and p.nt == 419
    assert (p.nx, p.nz, p.nx_abc, p.nz_abc) == (56, 40, 80, 64)
    assert p.Nshot == 3 and p.Nrec == 19
    assert p.sx.tolist() == [22, 37, 52] and p.sz.tolist() == [14, 14, 14]
    assert p.rx.tolist() == [12 + 3 * i for i in range(19)] and set(p.rz.tolist()) == {14}
    # C++ driver: std::round in float32
    q = parameters(path, rounding="round")
    assert q.nt_data == 301 and q.nt == 419 and q.itlag == 118
    assert p.sigma == 2.0 and p.shift == 0.04 and p.window == 0.05 and p.freqs == [30.0]


def test_itlag_rounding_differs_for_fcut_24():
    """tlag/dt = 147.70 at fcut = 24 Hz, dt = 1 ms: round gives 148, int() 147."""
    r = prep.derived_sizes(1000.0, 500.0, 1.0, 10.0, 10.0, 0.001, 24.0, "round")
    t = prep.derived_sizes(1000.0, 500.0, 1.0, 10.0, 10.0, 0.001, 24.0, "trunc")
    assert (r["itlag"], t["itlag"]) == (148, 147)
    assert r["nt"] == 148 + 1001 and t["nt"] == 147 + 1001


def test_reciprocity_and_mirror(tmp_path):
    path, _ = write_survey(str(tmp_path), nshots=2, extra={"reciprocity": True})
    p = parameters(path)
    rx, rz = read_geometry_csv(str(tmp_path / "rec.csv"))
    sx, sz = read_geometry_csv(str(tmp_path / "src.csv"))
    assert p.Nshot == len(rx) and p.Nrec == len(sx)
    np.testing.assert_array_equal(p.shot_x, rx)
    np.testing.assert_array_equal(p.rec_x, sx)
    path, _ = write_survey(str(tmp_path), nshots=2, extra={"mirror": True, "idx_water": 3})
    m = parameters(path)
    path, _ = write_survey(str(tmp_path), nshots=2)
    n = parameters(path)
    np.testing.assert_array_equal(m.sz, n.sz - 3)
    np.testing.assert_array_equal(m.rz, n.rz - 3)


def test_geometry_csv_without_header_columns(tmp_path):
    f = str(tmp_path / "g.csv")
    with open(f, "w") as h:
        h.write("0,10.0,5.0\n1,20.0,6.5\n")
    x, z = read_geometry_csv(f)
    assert x.tolist() == [10.0, 20.0] and z.tolist() == [5.0, 6.5]


@pytest.mark.parametrize("exe", ["run_modeling", "run_migration"])
def test_cli_parses_the_survey_as_the_cpp_driver(tmp_path, exe):
    """The C++ CLIs' Parameters.json / CSV reader agrees with the std::round
    rules (cpp/Survey.cpp:101-108, 150-158); ``--check`` does no GPU work."""
    binary = os.path.join(BIN, exe)
    if not os.path.exists(binary):
        pytest.skip("CLI not built (python -m pseudoacousticmodeling_amd.build)")
    path, _ = write_survey(str(tmp_path), phys=1, nshots=3, fcut=24.0, extra={"mirror": False})
    out = subprocess.run([binary, path, "--check"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.strip().splitlines()
    head = dict(zip(lines[0].split()[0::2], lines[0].split()[1::2]))
    q = parameters(path, rounding="round")
    assert int(head["nx"]) == q.nx and int(head["nz"]) == q.nz and int(head["nt"]) == q.nt
    assert int(head["itlag"]) == q.itlag == 148 and int(head["nt_data"]) == q.nt_data
    assert int(head["Nshot"]) == q.Nshot and int(head["Nrec"]) == q.Nrec
    assert head["fcut"] == "24.0"  # std::fixed << setprecision(1) (cpp/Modeling.cu:398-400)
    src = [tuple(map(int, l.split()[1:])) for l in lines if l.startswith("src")]
    rec = [tuple(map(int, l.split()[1:])) for l in lines if l.startswith("rec")]
    assert src == list(zip(q.sz.tolist(), q.sx.tolist()))
    assert rec == list(zip(q.rz.tolist(), q.rx.tolist()))


def test_cli_reports_a_missing_file(tmp_path):
    binary = os.path.join(BIN, "run_modeling")
    if not os.path.exists(binary):
        pytest.skip("CLI not built")
    out = subprocess.run([binary, str(tmp_path / "absent.json"), "--check"], capture_output=True, text=True,
                         timeout=60)
    assert out.returncode == 1 and "could not open" in out.stderr


# ---------------------------------------------------------------- preprocessing
def _gathers(seed=3, nt=300, nrec=19):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((nt, nrec)).astype(np.float32)


def test_mute_matches_oracle_and_its_closed_form():
    d = _gathers()
    rx = np.arange(0.0, 550.0, 30.0)
    rz = np.full(rx.shape, 20.0)
    a = prep.mute(d, rx, rz, 250.0, 20.0, 0.001, 0.04, 0.05, 1500.0)
    b = orc_mod.mute(d, rx, rz, 250.0, 20.0, 0.001, 0.04, 0.05, 1500.0)
    np.testing.assert_allclose(a, b, rtol=1e-6, atol=0)
    # closed form per trace: zero before t1, 1 after t2, linear in between
    for r in (0, 9, 18):
        t1 = abs(rx[r] - 250.0) / 1500.0 + 0.04
        t2 = t1 + 0.05
        t = np.arange(300) * 0.001
        assert np.all(a[t < t1 - 1e-9, r] == 0)
        np.testing.assert_array_equal(a[t >= t2 + 1e-9, r], d[t >= t2 + 1e-9, r])
        mid = (t > t1 + 1e-9) & (t < t2 - 1e-9)
        np.testing.assert_allclose(a[mid, r], d[mid, r] * (t[mid] - t1) / 0.05, rtol=1e-5)
    # the float32 (C++) form agrees to float32 rounding
    c = orc_mod.mute_cpp(d, rx, rz, 250.0, 20.0, 0.001, 0.04, 0.05, 1500.0)
    assert np.max(np.abs(c - b)) <= 1e-4 * np.max(np.abs(d))


def test_python_smoothing_matches_oracle_and_scipy():
    from scipy.ndimage import correlate

    rng = np.random.default_rng(11)
    vp = (2000.0 + 500.0 * rng.random((40, 56))).astype(np.float32)
    vp[:5] = 1500.0
    vp[:8, 30:] = 1500.0
    water = np.abs(vp - vp.min()) < 1e-3
    a = prep.smooth_model(vp, 2.0, water)
    b = orc_mod.smooth_model_py(vp, 2.0, water)
    np.testing.assert_allclose(a, b, rtol=2e-6)
    # interior: the normalised 2-D Gaussian correlation of the slowness
    k = orc_mod.gaussian_filter2d(2.0).astype(np.float64)
    h = k.shape[0] // 2
    ref = 1.0 / (correlate(1.0 / vp.astype(np.float64), k / k.sum(), mode="nearest"))
    inner = (slice(h, 40 - h), slice(h, 56 - h))
    dry = ~water[inner]
    np.testing.assert_allclose(b[inner][dry], ref[inner][dry], rtol=1e-5)
    # water column restored from the top down to the first dry cell
    assert np.all(b[:5] == 1500.0) and np.all(b[:8, 30:] == 1500.0)
    # border rows / columns replicate the first / last smoothed ones
    np.testing.assert_array_equal(b[h + 1:-h, 0], b[h + 1:-h, h])
    # a constant model is a fixed point
    c = np.full((30, 30), 2500.0, dtype=np.float32)
    np.testing.assert_allclose(orc_mod.smooth_model_py(c, 2.0, np.zeros_like(c, bool)), c, rtol=1e-6)


def test_cpp_smoothing_is_a_separable_clamped_gaussian():
    """cpp/Migration.cu:147-236 equals scipy's gaussian_filter1d along x then
    z with mode='nearest' and radius ceil(6 sigma + 1) // 2."""
    from scipy.ndimage import gaussian_filter1d

    rng = np.random.default_rng(12)
    vp = (2000.0 + 500.0 * rng.random((40, 56))).astype(np.float32)
    vp[:4] = 1500.0
    mask = np.abs(vp - vp.min()) < 1e-3
    for sigma in (1.5, 2.0, 3.0):
        h = int(np.ceil(6 * sigma + 1))
        h = (h + 1 if h % 2 == 0 else h) // 2
        s = 1.0 / vp.astype(np.float64)
        ref = gaussian_filter1d(gaussian_filter1d(s, sigma, axis=1, mode="nearest", radius=h), sigma, axis=0,
                                mode="nearest", radius=h)
        ref = np.where(mask, s, ref)
        out = orc_mod.smooth_model_cpp(vp, sigma, mask)
        np.testing.assert_allclose(1.0 / out, ref, rtol=2e-6)
        assert np.all(out[:4] == 1500.0)
        p = orc_mod.smooth_model_cpp(vp / 1e4, sigma, mask, parameter=True)
        np.testing.assert_allclose(p, np.where(mask, vp / 1e4, 1e-4 * gaussian_filter1d(
            gaussian_filter1d(vp.astype(np.float64), sigma, axis=1, mode="nearest", radius=h), sigma, axis=0,
            mode="nearest", radius=h)), rtol=2e-6)


# ---------------------------------------------------------------- L-BFGS
def test_two_loop_single_pair_is_the_bfgs_inverse_update():
    rng = np.random.default_rng(4)
    n = 7
    s = rng.standard_normal(n)
    y = s + 0.3 * rng.standard_normal(n)
    assert s @ y > 0
    gamma = (s @ y) / (y @ y)
    rho = 1.0 / (s @ y)
    I = np.eye(n)
    H = (I - rho * np.outer(s, y)) @ (gamma * I) @ (I - rho * np.outer(y, s)) + rho * np.outer(s, s)
    g = rng.standard_normal(n)
    np.testing.assert_allclose(two_loop_recursion(g, [s], [y]), H @ g, rtol=1e-12)
    # no pairs: the identity
    np.testing.assert_array_equal(two_loop_recursion(g, [], []), g)


def test_two_loop_satisfies_the_last_secant_condition():
    rng = np.random.default_rng(5)
    n = 9
    M = rng.standard_normal((n, n))
    A = M @ M.T + n * np.eye(n)
    S = [rng.standard_normal(n) for _ in range(5)]
    Y = [A @ s for s in S]
    np.testing.assert_allclose(two_loop_recursion(Y[-1], S, Y), S[-1], rtol=1e-10)


def test_two_loop_with_armijo_solves_rosenbrock():
    """src/optimazation.py:142-162: from (-2, -1.5), 30 L-BFGS iterations with
    Armijo backtracking (c1 = 1e-4, alpha = 1 halved) and the FWI driver's
    memory of 8 pairs (src/Inversion.py:391-393) reach (1, 1)."""
    f = lambda x: 100.0 * (x[1] - x[0] ** 2) ** 2 + (x[0] - 1) ** 2  # noqa: E731
    g = lambda x: np.array([-400.0 * x[0] * (x[1] - x[0] ** 2) + 2 * (x[0] - 1), 200.0 * (x[1] - x[0] ** 2)])  # noqa
    x = np.array([-2.0, -1.5])
    S, Y = [], []
    for _ in range(30):
        gr = g(x)
        if np.linalg.norm(gr) < 1e-12:
            break
        p = -two_loop_recursion(gr, S, Y)
        a = 1.0
        for _ in range(30):
            if f(x + a * p) <= f(x) + 1e-4 * a * (gr @ p):
                break
            a *= 0.5
        xn = x + a * p
        S.append(xn - x)
        Y.append(g(xn) - gr)
        if len(S) > 8:
            S.pop(0)
            Y.pop(0)
        x = xn
    np.testing.assert_allclose(x, [1.0, 1.0], atol=1e-6)


def test_write_csv_roundtrip(tmp_path):
    f = str(tmp_path / "r.csv")
    write_csv(f, [0.0, 12.5], [10.0, 10.0])
    x, z = read_geometry_csv(f)
    assert x.tolist() == [0.0, 12.5] and z.tolist() == [10.0, 10.0]
