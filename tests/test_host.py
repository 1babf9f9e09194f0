"""CPU tests of the host side: the C ABI library loads and exports every entry
point include/pam.h declares, and the host set-up (derived sizes, wavelets,
absorbing constants, grid indices) agrees with closed forms and with the
oracle's restatements."""
import ctypes
import os
import re

import numpy as np
import pytest

from oracle import oracle as orc
from pseudoacousticmodeling_amd import _lib, prep

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "pam.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(pam_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_entry_point():
    lib = _lib.load()
    names = _declared()
    assert len(names) >= 18
    for n in names:
        assert hasattr(lib, n), f"libpam.so does not export {n}"
    assert sorted(_lib.EXPORTS) == names


def test_version_and_null_context_errors():
    lib = _lib.load()
    assert b"gfx950" in lib.pam_version()
    assert lib.pam_destroy(None) == 0
    assert lib.pam_forward(None, 0, 1, None) == -1
    assert lib.pam_set_option(None, 1, 1) == -1


def test_context_free_calls_report_through_the_null_context():
    """The preprocessing entry points take no context; their argument checks
    run before any device call, and pam_last_error(NULL) explains them."""
    lib = _lib.load()
    assert lib.pam_smooth_model(None, None, None, 10, 10, 1.0, 0, 0, None) == -1
    assert b"pam_smooth_model" in lib.pam_last_error(None)
    D = ctypes.c_double
    x = (D * 2)(0.0, 1.0)
    assert lib.pam_mute(None, 1, 10, 2, x, x, x, x, 0.001, 0.0, 0.01, 0.0, 0, None) == -1
    assert b"pam_mute" in lib.pam_last_error(None)
    assert lib.pam_smooth_model(None, None, None, 10, 10, 1.0, 0, 7, None) == -1


def test_derived_sizes_round_and_trunc():
    # the 32-shot acoustic + CPML survey (fcut = 24, dt = 1 ms): itlag differs between the two drivers
    r = prep.derived_sizes(12000.0, 3000.0, 3.0, 10.0, 10.0, 1e-3, 24.0, "round")
    t = prep.derived_sizes(12000.0, 3000.0, 3.0, 10.0, 10.0, 1e-3, 24.0, "trunc")
    assert (r["itlag"], t["itlag"]) == (148, 147)
    assert r["nx"] == t["nx"] == 1201 and r["nz"] == t["nz"] == 301
    assert r["nt_data"] == t["nt_data"] == 3001
    # the one-shot two-layer acoustic survey: fcut = 30 -> tlag = 0.11816 s
    c1 = prep.derived_sizes(4000.0, 2000.0, 2.0, 10.0, 10.0, 1e-3, 30.0, "round")
    assert (c1["itlag"], c1["nt_data"], c1["nx"], c1["nz"], c1["nt"]) == (118, 2001, 401, 201, 2119)
    assert prep.derived_sizes(4000.0, 2000.0, 2.0, 10.0, 10.0, 1e-3, 30.0, "round") == \
        orc.derived_sizes(4000.0, 2000.0, 2.0, 10.0, 10.0, 1e-3, 30.0, "round")


def test_grid_index_rounding():
    assert list(prep.grid_index([14.9, 15.0, 25.0, 0.0], 10.0, 5, "round")) == [6, 7, 8, 5]
    assert list(prep.grid_index([14.9, 15.0, 25.0, 0.0], 10.0, 5, "trunc")) == [6, 6, 7, 5]


def test_ricker_closed_form_and_peak_frequency():
    nt, dt, fcut, dx = 2001, 1e-3, 30.0, 10.0
    tlag = 2 * np.sqrt(np.pi) / fcut
    w = prep.ricker_cpp(nt, dt, fcut, np.float32(tlag), dx, dx).astype(np.float64)
    assert np.array_equal(w, orc.ricker_cpp(nt, dt, fcut, np.float32(tlag), dx, dx))
    # standard Ricker with peak fp = fcut/3: (1 - 2 pi^2 fp^2 t^2) exp(-pi^2 fp^2 t^2) / (dx dz)
    fp = fcut / 3
    t = np.arange(nt) * dt - tlag
    ref = (1 - 2 * np.pi ** 2 * fp ** 2 * t ** 2) * np.exp(-np.pi ** 2 * fp ** 2 * t ** 2) / (dx * dx)
    assert np.max(np.abs(w - ref)) < 1e-6 * np.max(np.abs(ref))
    spec = np.abs(np.fft.rfft(w, 1 << 16))
    f = np.fft.rfftfreq(1 << 16, dt)
    assert abs(f[np.argmax(spec)] - fp) < 0.05


def test_python_ricker_half_derivative():
    nt, dt, fcut = 1024, 1e-3, 30.0
    tlag = 2 * np.sqrt(np.pi) / fcut
    w = prep.ricker_py(nt, dt, fcut, tlag, 1.0, 1.0).astype(np.float64)
    assert np.allclose(w, orc.ricker_py(nt, dt, fcut, tlag, 1.0, 1.0), rtol=0, atol=1e-7)
    # applying the same filter twice is the time derivative (sqrt(i w)^2 = i w)
    omega = 2 * np.pi * np.fft.fftfreq(nt, dt)
    twice = np.real(np.fft.ifft(np.fft.fft(w) * np.sqrt(1j * omega)))
    plain = prep.ricker_cpp(nt, dt, fcut, np.float32(tlag), 1.0, 1.0).astype(np.float64)
    deriv = np.real(np.fft.ifft(np.fft.fft(plain) * 1j * omega))
    assert np.max(np.abs(twice - deriv)) < 2e-3 * np.max(np.abs(deriv))


def test_cerjan_and_cpml_constants():
    A = prep.cerjan_vector(50)
    assert np.allclose(A, orc.cerjan_vector(50), rtol=1e-6)
    assert np.allclose(prep.cerjan_vector(50, 0.0045), orc.cerjan_vector(50, 0.0045), rtol=1e-6)
    d0, fp = prep.dampening_const(50, 30.0)
    assert abs(d0 - 3 * 5 * np.log(1e3)) < 1e-9 and fp == 10.0  # R = 1e-3^5
    d0b, _ = prep.dampening_const(250, 30.0)
    assert abs(d0b - 3 * 15 * np.log(1e3)) < 1e-9


def test_stability_check():
    s = prep.stability_acoustic(np.array([1500.0, 2500.0]), 30.0, 10.0, 10.0, 1e-3)
    assert abs(s["dx_lim"] - 1500 / 30 / 4.28) < 1e-12
    assert abs(s["dt_lim"] - s["dx_lim"] / (4 * 2500)) < 1e-15
    assert s["ok"] is True
    assert prep.stability_acoustic(np.array([1500.0, 2500.0]), 30.0, 12.0, 10.0, 1e-3)["ok"] is False


@pytest.mark.parametrize("kind", ["f64", "f32", "mixed"])
def test_oracle_builds_load(kind):
    o = orc.Oracle(kind)
    assert hasattr(o.lib, "or_forward_shot")


def test_create_rejects_a_plane_of_4_gib():
    """The kernels address one shot's field with a 32-bit byte offset from a
    per-shot base: pam_create refuses a padded plane of 4 GiB or more before
    any device call, with a message that says so (no GPU needed)."""
    lib = _lib.load()
    ctx = ctypes.c_void_p()
    # 32768 x 32768 padded cells x 4 B = 4 GiB exactly
    g = _lib.PamGrid(32768 - 100, 32768 - 100, 50, 10.0, 10.0, 1e-3, 100, 0)
    p = _lib.PamAbcParams(0.0, 0.0, 0.0)
    rc = lib.pam_create(ctypes.byref(g), 0, 0, ctypes.byref(p), None, 1, ctypes.byref(ctx))
    assert rc == -1
    assert b"4 GiB" in lib.pam_last_error(ctx)
    assert lib.pam_destroy(ctx) == 0
    # one row of pitch less is accepted by the size check (the call then goes on
    # to allocate, which this test does not reach): the bound is exact
    nzp, P = 32768, 32768
    assert nzp * P * 4 == 1 << 32 and (nzp - 1) * P * 4 < 1 << 32
