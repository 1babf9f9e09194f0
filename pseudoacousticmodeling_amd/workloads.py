"""Seeded synthetic workloads at the benchmark sizes.

Each builder returns a dict with the grid, the interior models (float32
[nz][nx]; theta in radians), the wavelet, and the source / receiver cell
indices on the padded grid.  The reference ships no model or geometry files
(its .bin and .csv inputs are git-ignored), so every workload here is built
from a fixed seed.
"""
from __future__ import annotations

import numpy as np

from . import prep

SEED0 = 20261203


def _gauss(f, sigma_cells):
    from scipy.ndimage import gaussian_filter

    return gaussian_filter(np.asarray(f, dtype=np.float64), sigma_cells, mode="nearest")


def _rescale(f, lo, hi):
    f = np.asarray(f, dtype=np.float64)
    span = max(float(f.max() - f.min()), 1e-30)
    return lo + (hi - lo) * (f - f.min()) / span


def vti_cpml_64(nshots=64, nt_data=5001, seed=SEED0 + 2):
    """Pseudo-acoustic VTI with a 50-cell CPML: nz=351, nx=1701, dx=dz=12.5 m,
    dt=1 ms, 5001 recorded samples; vp a linear gradient 1500->4500 m/s times
    (1 + 0.1 n), n Gaussian noise smoothed with sigma = 200 m and scaled to
    [-1, 1]; epsilon smoothed uniform in [0, 0.25]; delta in [0, min(eps, 0.1)];
    64 surface shots every 250 m from x = 2750 m, Ricker 12 Hz peak
    (fcut = 36 Hz); 1701 receivers at every grid column, z = 12.5 m."""
    rng = np.random.default_rng(seed)
    nz, nx, N, h, dt, fcut = 351, 1701, 50, 12.5, 1e-3, 36.0
    z = np.arange(nz) * h
    base = 1500.0 + 3000.0 * z / z[-1]
    n = _gauss(rng.standard_normal((nz, nx)), 200.0 / h)
    n = n / max(np.abs(n).max(), 1e-30)
    vp = (base[:, None] * (1.0 + 0.1 * n)).astype(np.float32)
    eps = _rescale(_gauss(rng.uniform(0.0, 0.25, (nz, nx)), 8.0), 0.0, 0.25).astype(np.float32)
    dfrac = _rescale(_gauss(rng.uniform(0.0, 1.0, (nz, nx)), 8.0), 0.0, 1.0)
    delta = (dfrac * np.minimum(eps, 0.1)).astype(np.float32)
    tlag = 2.0 * np.sqrt(np.pi) / fcut
    itlag = int(round(tlag / dt))
    nt = itlag + nt_data
    src = prep.ricker_cpp(nt, dt, fcut, np.float32(tlag), h, h)
    sx = prep.grid_index(2750.0 + 250.0 * np.arange(nshots), h, N)
    sz = prep.grid_index(np.full(nshots, 12.5), h, N)
    rx = prep.grid_index(np.arange(nx) * h, h, N)
    rz = prep.grid_index(np.full(nx, 12.5), h, N)
    d0, fp = prep.dampening_const(N, fcut)
    return dict(name="vti_cpml_64", physics=1, abc=1, nz=nz, nx=nx, N=N, dz=h, dx=h, dt=dt, nt=nt,
                itlag=itlag, fcut=fcut, vp=vp, eps=eps, delta=delta, theta=None, src=src,
                sz=sz, sx=sx, rz=rz, rx=rx, cpml_d0=d0, cpml_f_pico=fp, cerjan_factor=0.0)


def acoustic_cerjan_1():
    """Plain acoustic forward modelling with a 50-cell Cerjan strip (factor
    0.0045): nz=201, nx=401, dx=dz=10 m; vp = 1500 m/s above z = 1000 m and
    2500 m/s below; dt = 1 ms, 2001 recorded samples; Ricker 10 Hz peak
    (fcut = 30 Hz) at (z=10 m, x=2000 m); 400 receivers every 10 m at z=10 m."""
    nz, nx, N, h, dt, fcut = 201, 401, 50, 10.0, 1e-3, 30.0
    vp = np.full((nz, nx), 1500.0, dtype=np.float32)
    vp[100:, :] = 2500.0
    sz_ = prep.derived_sizes(4000.0, 2000.0, 2.0, h, h, dt, fcut, "round")
    itlag, nt = sz_["itlag"], sz_["nt"]
    src = prep.ricker_cpp(nt, dt, fcut, np.float32(sz_["tlag"]), h, h)
    sx = prep.grid_index([2000.0], h, N)
    sz = prep.grid_index([10.0], h, N)
    rx = prep.grid_index(np.arange(400) * h, h, N)
    rz = prep.grid_index(np.full(400, 10.0), h, N)
    return dict(name="acoustic_cerjan_1", physics=0, abc=0, nz=nz, nx=nx, N=N, dz=h, dx=h, dt=dt, nt=nt,
                itlag=itlag, fcut=fcut, vp=vp, eps=None, delta=None, theta=None, src=src, sz=sz, sx=sx,
                rz=rz, rx=rx, cpml_d0=0.0, cpml_f_pico=0.0, cerjan_factor=0.0045)


def tti_cpml_32(nshots=32, nt_data=4001, seed=SEED0 + 3):
    """Tilted TI with a 50-cell CPML (TTI + CPML extends the reference, whose
    Python driver stops there): nz=401, nx=2001, dx=dz=10 m, dt=0.8 ms, 4001
    recorded samples (T = 3.2 s); seven layers under six interfaces dipping
    0, 4, .. 20 degrees, vp 1500..4000 m/s plus a +800 m/s Gaussian lens of
    radius 300 m; epsilon 0..0.2 and delta 0..0.08 rising layer by layer; theta
    the dip of the interface above (radians; the reference reads degrees from
    disk); 32 surface shots spread over the model, Ricker 15 Hz peak
    (fcut = 45 Hz); receivers at every grid column, z = 10 m."""
    rng = np.random.default_rng(seed)
    nz, nx, N, h, dt, fcut = 401, 2001, 50, 10.0, 0.8e-3, 45.0
    z = np.arange(nz)[:, None] * h
    x = np.arange(nx)[None, :] * h
    dips = np.deg2rad(np.arange(6) * 4.0)
    z0 = np.linspace(400.0, 3600.0, 6) + rng.uniform(-50.0, 50.0, 6)
    xc = 0.5 * (nx - 1) * h
    layer = np.zeros((nz, nx), dtype=np.int64)
    theta = np.zeros((nz, nx))
    for i in range(6):
        below = z >= z0[i] + np.tan(dips[i]) * (x - xc)
        layer += below
        theta = np.where(below, dips[i], theta)
    v_l = np.linspace(1500.0, 4000.0, 7)
    vp = v_l[layer]
    lz, lx = 0.55 * (nz - 1) * h, 0.6 * (nx - 1) * h
    vp = vp + 800.0 * np.exp(-((z - lz) ** 2 + (x - lx) ** 2) / 300.0 ** 2)
    eps = np.linspace(0.0, 0.2, 7)[layer]
    delta = np.linspace(0.0, 0.08, 7)[layer]
    tlag = 2.0 * np.sqrt(np.pi) / fcut
    itlag = int(round(tlag / dt))
    nt = itlag + nt_data
    src = prep.ricker_cpp(nt, dt, fcut, np.float32(tlag), h, h)
    xs = (np.arange(nshots) + 0.5) * (nx - 1) * h / nshots
    return dict(name="tti_cpml_32", physics=2, abc=1, nz=nz, nx=nx, N=N, dz=h, dx=h, dt=dt, nt=nt, itlag=itlag,
                fcut=fcut, vp=vp.astype(np.float32), eps=eps.astype(np.float32), delta=delta.astype(np.float32),
                theta=theta.astype(np.float32), src=src, sz=prep.grid_index(np.full(nshots, 10.0), h, N),
                sx=prep.grid_index(xs, h, N), rz=prep.grid_index(np.full(nx, 10.0), h, N),
                rx=prep.grid_index(np.arange(nx) * h, h, N), cpml_d0=prep.dampening_const(N, fcut)[0],
                cpml_f_pico=prep.dampening_const(N, fcut)[1], cerjan_factor=0.0)


def acoustic_cpml_32(nshots=32, nt_data=3001, seed=SEED0 + 4):
    """Acoustic with a 50-cell CPML: nz=301, nx=1201, dx=dz=10 m, dt=1 ms, 3001
    recorded samples; six flat layers 1500..4500 m/s cut by a normal fault (the
    block right of x = 6 km dropped by 150 m) and a -300 m/s Gaussian anomaly of
    radius 200 m; 32 surface shots spread over the model, Ricker 8 Hz peak
    (fcut = 24 Hz, itlag 148 with the C++ rounding); receivers at every grid
    column, z = 10 m."""
    nz, nx, N, h, dt, fcut = 301, 1201, 50, 10.0, 1e-3, 24.0
    z = np.arange(nz)[:, None] * h
    x = np.arange(nx)[None, :] * h
    zf = z - np.where(x > 6000.0, 150.0, 0.0)
    layer = np.clip((zf // 500.0).astype(np.int64), 0, 5)
    vp = np.linspace(1500.0, 4500.0, 6)[layer]
    vp = vp - 300.0 * np.exp(-((z - 1800.0) ** 2 + (x - 3000.0) ** 2) / 200.0 ** 2)
    sz_ = prep.derived_sizes(12000.0, 3000.0, 3.0, h, h, dt, fcut, "round")
    itlag = sz_["itlag"]
    nt = itlag + nt_data
    src = prep.ricker_cpp(nt, dt, fcut, np.float32(sz_["tlag"]), h, h)
    xs = (np.arange(nshots) + 0.5) * (nx - 1) * h / nshots
    d0, fp = prep.dampening_const(N, fcut)
    return dict(name="acoustic_cpml_32", physics=0, abc=1, nz=nz, nx=nx, N=N, dz=h, dx=h, dt=dt, nt=nt,
                itlag=itlag, fcut=fcut, vp=vp.astype(np.float32), eps=None, delta=None, theta=None, src=src,
                sz=prep.grid_index(np.full(nshots, 10.0), h, N), sx=prep.grid_index(xs, h, N),
                rz=prep.grid_index(np.full(nx, 10.0), h, N), rx=prep.grid_index(np.arange(nx) * h, h, N),
                cpml_d0=d0, cpml_f_pico=fp, cerjan_factor=0.0)


def uneven_vti_group(shots_per_model=4, nt_data=3001, seed=SEED0 + 5):
    """Sixteen VTI models with a 50-cell CPML, nz cycling over
    {201, 301, 401, 601} and nx over {801, 1201, 1601, 2401} (every pair once),
    dx=dz=10 m, dt=1 ms, 3001 recorded samples, Ricker 10 Hz peak
    (fcut = 30 Hz); vp a smooth seeded random field rescaled to 1500-4500 m/s,
    epsilon smoothed uniform in [0, 0.25], delta in [0, min(eps, 0.1)];
    ``shots_per_model`` surface shots spread over each model; receivers at
    every grid column, z = 10 m.  Returns a list of workload dicts."""
    rng = np.random.default_rng(seed)
    h, dt, fcut, N = 10.0, 1e-3, 30.0, 50
    tlag = 2.0 * np.sqrt(np.pi) / fcut
    itlag = int(round(tlag / dt))
    nt = itlag + nt_data
    src = prep.ricker_cpp(nt, dt, fcut, np.float32(tlag), h, h)
    d0, fp = prep.dampening_const(N, fcut)
    out = []
    for i, nz in enumerate((201, 301, 401, 601)):
        for j, nx in enumerate((801, 1201, 1601, 2401)):
            vp = _rescale(_gauss(rng.standard_normal((nz, nx)), 20.0), 1500.0, 4500.0).astype(np.float32)
            eps = _rescale(_gauss(rng.uniform(0.0, 0.25, (nz, nx)), 8.0), 0.0, 0.25).astype(np.float32)
            dfrac = _rescale(_gauss(rng.uniform(0.0, 1.0, (nz, nx)), 8.0), 0.0, 1.0)
            delta = (dfrac * np.minimum(eps, 0.1)).astype(np.float32)
            xs = (np.arange(shots_per_model) + 0.5) * (nx - 1) * h / shots_per_model
            out.append(dict(name=f"uneven_vti_{nz}x{nx}", physics=1, abc=1, nz=nz, nx=nx, N=N, dz=h, dx=h,
                            dt=dt, nt=nt, itlag=itlag, fcut=fcut, vp=vp, eps=eps, delta=delta, theta=None,
                            src=src, sz=prep.grid_index(np.full(shots_per_model, 10.0), h, N),
                            sx=prep.grid_index(xs, h, N), rx=prep.grid_index(np.arange(nx) * h, h, N),
                            rz=prep.grid_index(np.full(nx, 10.0), h, N), cpml_d0=d0, cpml_f_pico=fp,
                            cerjan_factor=0.0))
    return out


def updates(w, nshots=None):
    """Padded cells x time steps x shots."""
    ns = len(w["sx"]) if nshots is None else nshots
    return (w["nz"] + 2 * w["N"]) * (w["nx"] + 2 * w["N"]) * w["nt"] * ns
