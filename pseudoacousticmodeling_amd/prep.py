"""Host-side set-up of a modelling run: derived sizes, source wavelets and the
absorbing-boundary constants.

These run once per run on the host (numpy) and feed the C ABI:

* ``derived_sizes``      -- cpp/Survey.cpp:101-108 (``round``) and
  src/survey.py:37-47 (``trunc``);
* ``ricker_cpp``         -- cpp/Modeling.cu:84-100 (plain Ricker, float32);
* ``ricker_py``          -- src/utils.py:21-28 with the 1/(dx dz) scale of
  src/Modeling2D.py:27-31 (Ricker with a sqrt(i w) half-derivative filter);
* ``cerjan_vector``      -- cpp/Modeling.cu:111-120 / src/Modeling2D.py:258-264;
* ``dampening_const``    -- src/Modeling2D.py:110-126 (CPML d0 and f_pico);
* ``grid_index``         -- source/receiver cell indices, cpp/Survey.cpp:150-158
  (``round``) and src/survey.py:131-138 (``trunc``);
* ``stability_acoustic`` -- cpp/Modeling.cu:287-292 / src/Modeling2D.py:202-223;
* ``stability_anisotropic`` -- the VTI / TTI rule, cpp/Modeling.cu:184-327 /
  src/Modeling2D.py:128-256 (qP phase-velocity scan, 8th-order dt limit).
"""
from __future__ import annotations

import math

import numpy as np

PI_F32 = np.float32(3.14159265358979323846)


def _cround(v) -> int:
    """std::round: halves away from zero."""
    v = float(v)
    return int(math.floor(abs(v) + 0.5)) * (1 if v >= 0 else -1)


def derived_sizes(L, D, T, dx, dz, dt, fcut, rounding="round"):
    """tlag = 2 sqrt(pi)/fcut, itlag, nt_data, nx, nz and nt = itlag + nt_data.

    ``round`` follows the C++ driver (float32 arithmetic, std::round);
    ``trunc`` follows the Python driver (float64 arithmetic, int())."""
    if rounding == "round":
        f32 = np.float32
        tlag = f32(2.0) * np.sqrt(PI_F32, dtype=np.float32) / f32(fcut)
        itlag = _cround(f32(tlag) / f32(dt))
        nt_data = _cround(f32(T) / f32(dt)) + 1
        nx = _cround(f32(L) / f32(dx)) + 1
        nz = _cround(f32(D) / f32(dz)) + 1
    elif rounding == "trunc":
        tlag = 2.0 * np.sqrt(np.pi) / fcut
        itlag = int(tlag / dt)
        nt_data = int(T / dt) + 1
        nx = int(L / dx) + 1
        nz = int(D / dz) + 1
    else:
        raise ValueError("rounding must be 'round' or 'trunc'")
    return dict(tlag=float(tlag), itlag=int(itlag), nt_data=int(nt_data), nx=int(nx), nz=int(nz),
                nt=int(itlag + nt_data))


def grid_index(coord, h, N, rounding="round"):
    """Padded-grid cell index of a coordinate: round(c/h)+N or int(c/h)+N."""
    c = np.asarray(coord, dtype=np.float64)
    if rounding == "round":
        q = np.asarray(c, dtype=np.float32) / np.float32(h)
        return (np.sign(q) * np.floor(np.abs(q) + 0.5)).astype(np.int32) + np.int32(N)
    return (c / h).astype(np.int32) + np.int32(N)


def ricker_cpp(nt, dt, fcut, tlag, dx, dz):
    """(1 - 2 arg) exp(-arg) / (dx dz), arg = pi^3 fc^2 td^2, fc = fcut/(3 sqrt(pi)),
    td = n dt - tlag; float32 arithmetic.  The peak frequency is fcut/3."""
    f32 = np.float32
    scale = f32(1.0) / (f32(dx) * f32(dz))
    fc = f32(fcut) / (f32(3.0) * np.sqrt(PI_F32, dtype=np.float32))
    n = np.arange(nt, dtype=np.float32)
    td = n * f32(dt) - f32(tlag)
    arg = PI_F32 * PI_F32 * PI_F32 * fc * fc * td * td
    return ((f32(1.0) - f32(2.0) * arg) * np.exp(-arg) * scale).astype(np.float32)


def ricker_py(nt, dt, fcut, tlag, dx, dz):
    """Ricker followed by the sqrt(i w) half-derivative filter applied with an
    FFT, scaled by 1/(dx dz); float64, returned as float32 (the cast the
    reference makes when it uploads the wavelet)."""
    pi = np.pi
    f = fcut / (3 * np.sqrt(pi))
    td = np.arange(nt) * dt - tlag
    src = (1 - 2 * pi * (pi * f * td) ** 2) * np.exp(-pi * (pi * f * td) ** 2)
    w = 2.0 * pi * np.fft.fftfreq(nt, dt)
    src = np.real(np.fft.ifft(np.fft.fft(src) * np.sqrt(1j * w)))
    return (src / (dx * dz)).astype(np.float32)


def cerjan_vector(N, factor=0.0):
    """A[i] = exp(-fb^2), i < N; fb = (N-i)/(1.4142*3N) for factor <= 0 (the
    reference's taper), else fb = factor*(N-i).  float32 as on the device."""
    i = np.arange(N, dtype=np.float32)
    if factor is None or factor <= 0:
        fb = (np.float32(N) - i) / (np.float32(1.4142) * (np.float32(3.0) * np.float32(N)))
    else:
        fb = np.float32(factor) * (np.float32(N) - i)
    return np.exp(-fb * fb).astype(np.float32)


def dampening_const(N, fcut):
    """CPML constants: R = 1e-3^(N/10) (1e-3^15 for N >= 200), d0 = -(M+1) ln R
    with M = 2, f_pico = fcut/3."""
    R = 1e-3 ** (N / 10)
    if N >= 200:
        R = 1e-3 ** (150 / 10)
    return float(-(2 + 1) * np.log(R)), float(fcut / 3)


def stability_acoustic(vp, fcut, dx, dz, dt):
    """lambda_min = vmin/fcut, dx_lim = lambda_min/4.28, dt_lim = dx_lim/(4 vmax)."""
    lam = float(np.min(vp)) / fcut
    dx_lim = lam / 4.28
    dt_lim = dx_lim / (4.0 * float(np.max(vp)))
    return dict(dx_lim=dx_lim, dz_lim=dx_lim, dt_lim=dt_lim,
                ok=bool(dx <= dx_lim and dz <= dx_lim and dt <= dt_lim))


# sum over the 8th-order second-derivative coefficients c_m (1 - (-1)^m), the
# stencil's gain at the Nyquist wavenumber (src/Modeling2D.py:144-149,
# cpp/Modeling.cu:203-208)
_COEFFS_8TH = (8.0 / 5.0, -1.0 / 5.0, 8.0 / 315.0, -1.0 / 560.0)


def stability_sum(coeffs=_COEFFS_8TH):
    return float(sum(c * (1.0 - (-1.0) ** m) for m, c in enumerate(coeffs, start=1)))


def _sd_unit(eps, dl, s2, c2, gate):
    """Sd on a unit plane wave with sin^2 = s2, cos^2 = c2 (relative to the
    symmetry axis): -2 (eps - delta) s2 c2 / ((1 + 2 eps) s2^2 + c2^2 +
    2 (1 + delta) s2 c2); with `gate` 0 where |den| <= 1e-12 (the C++ rule)."""
    num = -2.0 * (eps - dl) * s2 * c2
    den = (1.0 + 2.0 * eps) * (s2 * s2) + (c2 * c2) + 2.0 * (1.0 + dl) * s2 * c2
    if gate:
        return np.where(np.abs(den) > 1e-12, num / np.where(np.abs(den) > 1e-12, den, 1.0), 0.0).astype(num.dtype)
    return num / den


def stability_anisotropic(vp, eps, delta, fcut, dx, dz, dt, theta=None, rule="py"):
    """Dispersion and stability limits of the pseudo-acoustic VTI / TTI step:
    src/Modeling2D.py:128-256 (``rule='py'``, float64) and
    cpp/Modeling.cu:184-327 (``rule='cpp'``, float32, Sd gated at |den| <= 1e-12).

    * dt_lim = sqrt(2) / (sqrt(S) max(vp sqrt(A))), S = sum c_m (1 - (-1)^m) of the
      8th-order coefficients, A = (1 + 2 eps + Sk) / dx^2 + (1 + Sk) / dz^2 with
      Sk the Sd term at the grid's Nyquist wavenumbers (1/dx, 1/dz);
    * the qP phase velocity vp sqrt((1 + 2 eps) sin^2 + cos^2 + Sd) over 181
      angles in [0, pi/2] (VTI) or 361 angles phi in [0, pi] relative to the tilt,
      beta = phi - theta (TTI, theta in radians); its minimum gives
      dx_lim = dz_lim = vqp_min / fcut / 4.28.
    Returns vqp_min, vqp_max, eps_max, delta_max, dx_lim, dz_lim, dt_lim, ok."""
    f = np.float64 if rule == "py" else np.float32
    gate = rule == "cpp"
    vp = np.asarray(vp, dtype=f).ravel()
    eps = np.asarray(eps, dtype=f).ravel()
    dl = np.asarray(delta, dtype=f).ravel()
    th = None if theta is None else np.asarray(theta, dtype=f).ravel()
    one, two = f(1.0), f(2.0)
    idx2 = (one / f(dx)) * (one / f(dx))
    idz2 = (one / f(dz)) * (one / f(dz))
    # Sk at the Nyquist wavenumbers: the unit-wave formula with s2, c2 -> 1/dx^2, 1/dz^2
    Sk = _sd_unit(eps, dl, idx2, idz2, gate).astype(f)
    A = ((one + two * eps) + Sk) * idx2 + (one + Sk) * idz2
    soma = f(stability_sum())
    dt_lim = float(f(np.sqrt(f(2.0))) / (np.sqrt(soma) * np.max(vp * np.sqrt(A))))
    if th is None:
        angles = np.linspace(0.0, 0.5 * np.pi, 181)
    else:
        angles = np.linspace(0.0, np.pi, 361)
    vmin, vmax = np.inf, 0.0
    for phi in angles:
        beta = f(phi) if th is None else f(phi) - th
        sn, cs = np.sin(beta), np.cos(beta)
        s2, c2 = sn * sn, cs * cs
        factor = (one + two * eps) * s2 + c2 + _sd_unit(eps, dl, s2, c2, gate).astype(f)
        v = vp * np.sqrt(factor)
        vmin = min(vmin, float(np.min(v)))
        vmax = max(vmax, float(np.max(v)))
    lam = vmin / fcut
    dx_lim = lam / 4.28
    return dict(vqp_min=vmin, vqp_max=vmax, eps_max=float(np.max(eps)), delta_max=float(np.max(dl)),
                dx_lim=dx_lim, dz_lim=dx_lim, dt_lim=dt_lim,
                ok=bool(dx <= dx_lim and dz <= dx_lim and dt <= dt_lim))


def expand_model(m, N):
    """Edge-replicating pad by N cells on every side, corners taking the corner
    value: src/Modeling2D.py:38-54 / cpp/Modeling.cu:130-182 (the device path
    pads on the GPU; this host copy serves drivers that write padded files)."""
    return np.pad(np.asarray(m, dtype=np.float32), N, mode="edge")


def mute(seis, rec_x, rec_z, shot_x, shot_z, dt, shift, window, v0):
    """Linear top mute of one shot gather [nt][nrec]: zero before
    t1 = dist/v0 + shift, a linear ramp to t2 = t1 + window, unchanged after
    (src/utils.py:30-49, cpp/Migration.cu:86-115)."""
    seis = np.asarray(seis, dtype=np.float32)
    nt, nrec = seis.shape
    dist = np.sqrt((np.asarray(rec_z, dtype=np.float64) - shot_z) ** 2 +
                   (np.asarray(rec_x, dtype=np.float64) - shot_x) ** 2)
    t1 = dist / v0 + shift
    t2 = t1 + window
    t = (np.arange(nt) * dt)[:, None]
    ramp = (t >= t1[None, :]) & (t < t2[None, :])
    w = np.where(t >= t2[None, :], 1.0, np.where(ramp, (t - t1[None, :]) / (t2 - t1)[None, :], 0.0))
    return (w * seis).astype(np.float32)


def gaussian_filter2d(sigma):
    """Normalised 2-D Gaussian kernel of size ceil(6 sigma + 1) made odd
    (src/utils.py:51-74)."""
    ks = int(np.ceil(6.0 * sigma + 1))
    if ks % 2 == 0:
        ks += 1
    h = ks // 2
    x = (np.arange(ks) - h).astype(np.float64)
    val = (1.0 / (2.0 * np.pi * sigma * sigma)) * np.exp(-(x[:, None] ** 2 + x[None, :] ** 2) / (2.0 * sigma * sigma))
    return (val.astype(np.float32) / np.float32(val.sum())).astype(np.float32)


def smooth_model(f, sigma, water_mask, parameter=False):
    """Gaussian smoothing of slowness (or of a parameter when ``parameter``)
    over the cells a full kernel fits, edge replication of the border rows and
    columns, and restoration of the water column from the top of each column
    down to its first non-water cell (src/utils.py:76-129)."""
    f = np.asarray(f, dtype=np.float32)
    s = f.copy() if parameter else (np.float32(1.0) / f)
    s_old = s.copy()
    k = gaussian_filter2d(sigma)
    ks = k.shape[0]
    h = ks // 2
    nz, nx = s.shape
    if nz > 2 * h and nx > 2 * h:
        from numpy.lib.stride_tricks import sliding_window_view

        win = sliding_window_view(s_old, (ks, ks))
        acc = np.einsum("abij,ij->ab", win.astype(np.float64), k.astype(np.float64))
        s[h:nz - h, h:nx - h] = (acc / float(k.astype(np.float64).sum())).astype(np.float32)
    for z in range(h):
        s[z, :] = s[h, :]
        s[nz - 1 - z, :] = s[nz - 1 - h, :]
    for x in range(h):
        s[:, x] = s[:, h]
        s[:, nx - 1 - x] = s[:, nx - 1 - h]
    wm = np.asarray(water_mask, dtype=bool)
    # the water column: rows above the first non-water cell of each column
    first_dry = np.where(wm.all(axis=0), nz, np.argmin(wm, axis=0))
    rows = np.arange(nz)[:, None] < first_dry[None, :]
    s[rows] = s_old[rows]
    return s if parameter else (np.float32(1.0) / s).astype(np.float32)


def poisson_disk(x_max, z_max, radius, rng, k=30):
    """Bridson Poisson-disk samples in [0, x_max) x [0, z_max), no two closer
    than ``radius`` (src/Migration.py:333-380, with a seeded generator)."""
    cell = radius / np.sqrt(2.0)
    gw, gh = int(x_max / cell) + 1, int(z_max / cell) + 1
    grid = -np.ones((gw, gh), dtype=np.int64)
    pts = []

    def ok(p):
        gx, gz = int(p[0] / cell), int(p[1] / cell)
        for i in range(max(gx - 2, 0), min(gx + 3, gw)):
            for j in range(max(gz - 2, 0), min(gz + 3, gh)):
                n = grid[i, j]
                if n >= 0 and (pts[n][0] - p[0]) ** 2 + (pts[n][1] - p[1]) ** 2 < radius * radius:
                    return False
        return True

    p0 = (rng.uniform(0, x_max), rng.uniform(0, z_max))
    pts.append(p0)
    grid[int(p0[0] / cell), int(p0[1] / cell)] = 0
    active = [0]
    while active:
        a = int(rng.integers(len(active)))
        p = pts[active[a]]
        found = False
        for _ in range(k):
            r = rng.uniform(radius, 2 * radius)
            ang = rng.uniform(0, 2 * np.pi)
            q = (p[0] + r * np.cos(ang), p[1] + r * np.sin(ang))
            if 0 <= q[0] < x_max and 0 <= q[1] < z_max and ok(q):
                pts.append(q)
                grid[int(q[0] / cell), int(q[1] / cell)] = len(pts) - 1
                active.append(len(pts) - 1)
                found = True
        if not found:
            active.pop(a)
    return np.array(pts)


def random_boundary(vp_exp, N, dx, dz, dvel, ratio, rng):
    """Random-boundary velocity model of src/Migration.py:382-458 on a padded
    [nzp][nxp] model: every absorbing-layer cell drawn from a normal
    distribution around the velocity at the interior edge, its spread growing
    linearly (f1d) towards the outer edge and the draw limited to
    [dvel, v + dvel] (get_randomvalue, src/Migration.py:327-331); the model
    clipped to [vmin - dvel, vmax + dvel]; then Gaussian blobs (radius
    U(0.1, 1) ratio, amplitude +-U(0.5, 1) dvel) centred on Poisson-disk points
    in the outer half of the layer added over the grid, and clipped again.
    ``rng`` is a numpy Generator: the reference draws from the global
    generators, so realisations are comparable only statistically."""
    vp = np.array(vp_exp, dtype=np.float64)
    nzp, nxp = vp.shape
    nz, nx = nzp - 2 * N, nxp - 2 * N
    vmax, vmin = vp.max(), vp.min()
    f1d = np.linspace(0.0, 1.0, N)

    def rv(v, f):
        p = rng.normal(v, dvel * f)
        return np.where(p < dvel, dvel, np.where(p > dvel + v, dvel + v, p))

    ramp = f1d[::-1]  # spread at distance j+1 .. N from the interior: f1d[N-j-1] for column j
    # left / right columns of the interior rows
    vl = vp[N:N + nz, N][:, None]
    vr = vp[N:N + nz, nxp - N - 1][:, None]
    vp[N:N + nz, :N] = rv(np.repeat(vl, N, 1), np.repeat(ramp[None, :], nz, 0))
    vp[N:N + nz, nxp - N:] = rv(np.repeat(vr, N, 1), np.repeat(ramp[None, ::-1], nz, 0))
    # top / bottom rows of the interior columns
    vt = vp[N, N:N + nx][None, :]
    vb = vp[nzp - N - 1, N:N + nx][None, :]
    vp[:N, N:N + nx] = rv(np.repeat(vt, N, 0), np.repeat(ramp[:, None], nx, 1))
    vp[nzp - N:, N:N + nx] = rv(np.repeat(vb, N, 0), np.repeat(ramp[::-1, None], nx, 1))
    # corners: spread by the distance to the interior corner (the larger of the two)
    ii, jj = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
    dist = np.maximum(N - 1 - ii, N - 1 - jj)  # 0 next to the interior
    fc = f1d[dist]
    for (zs, xs, vc) in ((slice(0, N), slice(0, N), vp[N, N]), (slice(0, N), slice(nxp - N, nxp), vp[N, nxp - N - 1]),
                         (slice(nzp - N, nzp), slice(0, N), vp[nzp - N - 1, N]),
                         (slice(nzp - N, nzp), slice(nxp - N, nxp), vp[nzp - N - 1, nxp - N - 1])):
        f = fc
        if zs.start != 0:
            f = f[::-1, :]
        if xs.start != 0:
            f = f[:, ::-1]
        vp[zs, xs] = rv(np.full((N, N), vc), f)
    vp = np.clip(vp, vmin - dvel, vmax + dvel)
    # blobs at Poisson-disk points in the outer half of the layer
    L_abc, D_abc = (nxp - 1) * dx, (nzp - 1) * dz
    pts = poisson_disk(L_abc, D_abc, ratio, rng)
    nab_x, nab_z = N * dx, N * dz
    L, D = (nx - 1) * dx, (nz - 1) * dz
    keep = ((pts[:, 0] < 0.5 * nab_x) | (pts[:, 0] > L + nab_x + 0.5 * nab_x) |
            (pts[:, 1] < 0.5 * nab_z) | (pts[:, 1] > D + nab_z + 0.5 * nab_z))
    x = np.arange(nxp)[None, :] * dx
    z = np.arange(nzp)[:, None] * dz
    for xc, zc in pts[keep]:
        r = rng.uniform(0.1 * ratio, ratio)
        A = rng.uniform(0.5 * dvel, dvel)
        sgn = rng.choice([-1.0, 1.0])
        vp = vp + sgn * A * np.exp(-0.5 * (((x - xc) / r) ** 2 + ((z - zc) / r) ** 2))
    vp = np.clip(vp, vmin - dvel, vmax + dvel)
    return vp.astype(np.float32)
