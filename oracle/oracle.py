"""CPU oracle for the 2D finite-difference engine -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg
import this module, and only as the checker.  The product package never does.

Two layers:

* ctypes bindings to ``pam_oracle.c`` built three ways (``f64``: float64
  storage and arithmetic; ``f32``: float32 storage and arithmetic without FMA
  contraction; ``mixed``: float32 storage, float64 arithmetic, OpenMP over rows
  -- the Numba kernels' numerics and loop structure, used as the timed CPU
  baseline);
* numpy restatements of the small host-side pieces (wavelet, mute, smoothing,
  stability check, derived sizes) and whole-run drivers (forward shot, RTM and
  FWI gradient, on-the-fly strategy) that follow the reference's Python loops.

Every function cites the reference file:line it restates.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS: dict = {}


def build() -> None:
    """Compile the three oracle libraries (gcc, make)."""
    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _lib(kind: str):
    if kind not in _LIBS:
        path = os.path.join(_HERE, f"liboracle_{kind}.so")
        if not os.path.exists(path):
            build()
        lib = ctypes.CDLL(path)
        _LIBS[kind] = lib
    return _LIBS[kind]


_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


# --------------------------------------------------------------------------
# grid description and derived sizes
# --------------------------------------------------------------------------
@dataclass
class Grid:
    nz: int
    nx: int
    N: int
    dz: float
    dx: float
    dt: float
    nt: int
    itlag: int = 0

    @property
    def nzp(self):
        return self.nz + 2 * self.N

    @property
    def nxp(self):
        return self.nx + 2 * self.N

    @property
    def nt_data(self):
        return self.nt - self.itlag


def derived_sizes(L, D, T, dx, dz, dt, fcut, rounding="round"):
    """Derived sizes of cpp/Survey.cpp:101-108 (``round``) and
    src/survey.py:37-47 (``trunc``): tlag = 2 sqrt(pi)/fcut, itlag, nt_data,
    nx, nz, nt = itlag + nt_data."""
    if rounding == "round":
        f32 = np.float32
        cround = lambda v: int(np.floor(abs(float(v)) + 0.5)) * (1 if v >= 0 else -1)  # noqa: E731
        tlag = f32(2.0) * np.sqrt(f32(3.14159265358979323846)) / f32(fcut)
        itlag = cround(f32(tlag) / f32(dt))
        nt_data = cround(f32(T) / f32(dt)) + 1
        nx = cround(f32(L) / f32(dx)) + 1
        nz = cround(f32(D) / f32(dz)) + 1
    else:
        tlag = 2.0 * np.sqrt(np.pi) / fcut
        itlag = int(tlag / dt)
        nt_data = int(T / dt) + 1
        nx = int(L / dx) + 1
        nz = int(D / dz) + 1
    return dict(tlag=float(tlag), itlag=itlag, nt_data=nt_data, nx=nx, nz=nz, nt=itlag + nt_data)


# --------------------------------------------------------------------------
# numpy restatements of host-side pieces
# --------------------------------------------------------------------------
def ricker_cpp(nt, dt, fcut, tlag, dx, dz):
    """Plain Ricker of cpp/Modeling.cu:84-100, float32 arithmetic:
    (1-2*pi^3 f^2 td^2) exp(-pi^3 f^2 td^2) / (dx dz), f = fcut/(3 sqrt(pi))."""
    f32 = np.float32
    pi = f32(3.14159265358979323846)
    scale = f32(1.0) / (f32(dx) * f32(dz))
    fc = f32(fcut) / (f32(3.0) * np.sqrt(pi, dtype=np.float32))
    n = np.arange(nt, dtype=np.float32)
    td = n * f32(dt) - f32(tlag)
    arg = pi * pi * pi * fc * fc * td * td
    return ((f32(1.0) - f32(2.0) * arg) * np.exp(-arg) * scale).astype(np.float32)


def ricker_py(nt, dt, fcut, tlag, dx, dz):
    """Ricker of src/utils.py:21-28 with the sqrt(i w) half-derivative filter,
    scaled by 1/(dx dz) as src/Modeling2D.py:27-31; float64, cast to float32
    when uploaded (src/Modeling2D.py:427)."""
    pi = np.pi
    f = fcut / (3 * np.sqrt(pi))
    t = np.arange(nt) * dt
    td = t - tlag
    src = (1 - 2 * pi * (pi * f * td) ** 2) * np.exp(-pi * (pi * f * td) ** 2)
    w = 2.0 * pi * np.fft.fftfreq(nt, dt)
    src = np.real(np.fft.ifft(np.fft.fft(src) * np.sqrt(1j * w)))
    return (src / (dx * dz)).astype(np.float32)


def cerjan_vector(N, factor=None):
    """src/Modeling2D.py:258-264 (float64) -- factor None: the reference's
    fb = (N-i)/(1.4142*3N); otherwise fb = factor*(N-i)."""
    i = np.arange(N, dtype=np.float64)
    if factor is None or factor <= 0:
        fb = (N - i) / (1.4142 * 3.0 * N)
    else:
        fb = factor * (N - i)
    return np.exp(-fb * fb)


def dampening_const(N, fcut):
    """src/Modeling2D.py:110-126: R = 1e-3^(N/10) (N >= 200: 1e-3^15),
    d0 = -(M+1) ln R with M = 2, f_pico = fcut/3."""
    R = 1e-3 ** (N / 10)
    if N >= 200:
        R = 1e-3 ** (150 / 10)
    return -(2 + 1) * np.log(R), fcut / 3


def expand_model(m, N):
    """src/Modeling2D.py:38-54 (edge replication; corners take corner values)."""
    return np.pad(np.asarray(m, dtype=np.float32), N, mode="edge")


def mute(seis, rec_x, rec_z, sx, sz, dt, shift, window, v0):
    """Linear top mute, src/utils.py:30-49 (== cpp/Migration.cu:86-115)."""
    nt, nrec = seis.shape
    out = np.zeros_like(seis)
    t = np.arange(nt) * dt
    for r in range(nrec):
        dist = np.sqrt((rec_z[r] - sz) ** 2 + (rec_x[r] - sx) ** 2)
        t1 = dist / v0 + shift
        t2 = t1 + window
        ramp = (t >= t1) & (t < t2)
        out[ramp, r] = ((t[ramp] - t1) / (t2 - t1) * seis[ramp, r]).astype(seis.dtype)
        full = t >= t2
        out[full, r] = seis[full, r]
    return out


def gaussian_filter2d(sigma):
    """src/utils.py:51-74."""
    ks = int(np.ceil(6.0 * sigma + 1))
    if ks % 2 == 0:
        ks += 1
    h = ks // 2
    x = np.arange(ks) - h
    X, Z = np.meshgrid(x, x, indexing="ij")
    k = (1.0 / (2 * np.pi * sigma * sigma)) * np.exp(-(X * X + Z * Z) / (2 * sigma * sigma))
    k = k.astype(np.float32)
    total = np.float64(0.0)
    for v in (1.0 / (2 * np.pi * sigma * sigma)) * np.exp(-(X * X + Z * Z) / (2 * sigma * sigma)).ravel():
        total += v
    return (k / np.float32(total)).astype(np.float32)


def smooth_model_py(f, sigma, water_mask, parameter=False):
    """2-D Gaussian smoothing of slowness (or of a parameter) over the inner
    region, edge replication of the border, water column restored:
    src/utils.py:76-129."""
    f = np.asarray(f, dtype=np.float32)
    s = f.copy() if parameter else (np.float32(1.0) / f)
    s_old = s.copy()
    k = gaussian_filter2d(sigma)
    ks = k.shape[0]
    h = ks // 2
    nz, nx = s.shape
    from numpy.lib.stride_tricks import sliding_window_view

    if nz > 2 * h and nx > 2 * h:
        win = sliding_window_view(s_old.astype(np.float64), (ks, ks))
        conv = np.einsum("abij,ij->ab", win, k.astype(np.float64))
        total = float(np.sum(k.astype(np.float64)))
        s[h:nz - h, h:nx - h] = (conv / total).astype(np.float32)
    for z in range(h):
        s[z, :] = s[h, :]
        s[nz - 1 - z, :] = s[nz - 1 - h, :]
    for x in range(h):
        s[:, x] = s[:, h]
        s[:, nx - 1 - x] = s[:, nx - 1 - h]
    for x in range(nx):
        for z in range(nz):
            if water_mask[z, x]:
                s[z, x] = s_old[z, x]
            else:
                break
    return s if parameter else (np.float32(1.0) / s)


def mute_cpp(seis, rec_x, rec_z, sx, sz, dt, shift, window, v0):
    """Linear top mute of the C++ driver, cpp/Migration.cu:86-115: the same
    rule as ``mute`` with every quantity a float32 (sqrtf of the float
    offsets, t = it * dt in float32, the ramp factor multiplied in place)."""
    f32 = np.float32
    out = np.array(seis, dtype=np.float32, copy=True)
    nt, nrec = out.shape
    for r in range(nrec):
        ddz = f32(rec_z[r]) - f32(sz)
        ddx = f32(rec_x[r]) - f32(sx)
        dist = np.sqrt(f32(ddz * ddz + ddx * ddx), dtype=np.float32)
        t1 = f32(dist / f32(v0)) + f32(shift)
        t2 = f32(t1 + f32(window))
        for it in range(nt):
            t = f32(it) * f32(dt)
            if t < t1:
                out[it, r] = 0.0
            elif t < t2:
                out[it, r] = out[it, r] * f32((t - t1) / (t2 - t1))
    return out


def gaussian_kernel1d_cpp(sigma):
    """Normalised 1-D Gaussian of size ceil(6 sigma + 1) made odd, float32
    with a sequential float32 sum: cpp/Migration.cu:147-169."""
    f32 = np.float32
    ks = int(np.ceil(f32(6.0) * f32(sigma) + f32(1.0)))
    if ks % 2 == 0:
        ks += 1
    h = ks // 2
    k = np.zeros(ks, dtype=np.float32)
    total = f32(0.0)
    for i in range(ks):
        x = i - h
        k[i] = np.exp(f32(-(x * x)) / (f32(2.0) * f32(sigma) * f32(sigma)), dtype=np.float32)
        total = f32(total + k[i])
    return (k / total).astype(np.float32)


def smooth_model_cpp(f, sigma, mask, parameter=False):
    """Separable Gaussian smoothing of the C++ driver, cpp/Migration.cu:171-236:
    slowness (or the parameter) filtered along x then along z with clamped
    (edge-replicated) indices, every masked cell restored, slowness turned back
    into velocity.  float32 with float32 accumulation in tap order."""
    f = np.asarray(f, dtype=np.float32)
    old = f.copy() if parameter else (np.float32(1.0) / f).astype(np.float32)
    k = gaussian_kernel1d_cpp(sigma)
    h = k.shape[0] // 2
    nz, nx = old.shape
    xi = np.arange(nx)
    zi = np.arange(nz)
    tmp = np.zeros_like(old)
    for i in range(-h, h + 1):
        tmp = (tmp + k[i + h] * old[:, np.clip(xi + i, 0, nx - 1)]).astype(np.float32)
    out = np.zeros_like(old)
    for i in range(-h, h + 1):
        out = (out + k[i + h] * tmp[np.clip(zi + i, 0, nz - 1), :]).astype(np.float32)
    out = np.where(np.asarray(mask, dtype=bool), old, out).astype(np.float32)
    return out if parameter else (np.float32(1.0) / out).astype(np.float32)


def stability_acoustic(vp, fcut, dx, dz, dt):
    """src/Modeling2D.py:202-223 / cpp/Modeling.cu:287-292: lambda_min = vmin/fcut,
    dx_lim = lambda_min/4.28, dt_lim = dx_lim/(4 vmax)."""
    lam = np.min(vp) / fcut
    dx_lim = lam / 4.28
    dt_lim = dx_lim / (4.0 * np.max(vp))
    return dict(dx_lim=dx_lim, dz_lim=dx_lim, dt_lim=dt_lim,
                ok=bool(dx <= dx_lim and dz <= dx_lim and dt <= dt_lim))


# --------------------------------------------------------------------------
# ctypes layer
# --------------------------------------------------------------------------
class Oracle:
    """Per-kernel bindings.  ``kind``: 'f64' | 'f32' | 'mixed'."""

    def __init__(self, kind="f64"):
        self.kind = kind
        self.lib = _lib(kind)
        self.dtype = np.float64 if kind == "f64" else np.float32

    def arr(self, a):
        return np.ascontiguousarray(a, dtype=self.dtype)

    def zeros(self, *shape):
        return np.zeros(shape, dtype=self.dtype)

    def expand(self, m, N):
        m = np.ascontiguousarray(m, dtype=np.float32)
        nz, nx = m.shape
        out = self.zeros(nz + 2 * N, nx + 2 * N)
        self.lib.or_expand(_ptr(m), _I(nz), _I(nx), _I(N), _ptr(out))
        return out

    def cerjan(self, N, factor=0.0):
        A = self.zeros(N)
        self.lib.or_cerjan(_I(N), _D(factor if factor else 0.0), _ptr(A))
        return A

    def cpml_profiles(self, vp_exp, N, dx, dz, dt, d0, fpico):
        nzp, nxp = vp_exp.shape
        outs = [self.zeros(nzp, nxp) for _ in range(4)]
        self.lib.or_cpml_profiles(_ptr(self.arr(vp_exp)), _I(nzp), _I(nxp), _I(N), _D(dx), _D(dz),
                                  _D(dt), _D(d0), _D(fpico), *[_ptr(o) for o in outs])
        return outs

    # single kernels -------------------------------------------------------
    def update(self, phys, Uf, Uc, g: Grid, vp, eps=None, dl=None, th=None):
        L = self.lib
        if phys == 0:
            L.or_update_acoustic(_ptr(Uf), _ptr(Uc), _ptr(vp), _I(g.nzp), _I(g.nxp), _D(g.dz), _D(g.dx), _D(g.dt))
        elif phys == 1:
            L.or_update_vti(_ptr(Uf), _ptr(Uc), _ptr(vp), _ptr(eps), _ptr(dl), _I(g.nzp), _I(g.nxp),
                            _D(g.dz), _D(g.dx), _D(g.dt))
        else:
            L.or_update_tti(_ptr(Uf), _ptr(Uc), _ptr(vp), _ptr(eps), _ptr(dl), _ptr(th), _I(g.nzp),
                            _I(g.nxp), _D(g.dz), _D(g.dx), _D(g.dt))

    def operator(self, phys, Uc, g: Grid, vp, eps=None, dl=None, th=None):
        """The operator term V L(Uc) of the undamped update (zero on the
        outer 4-cell rim): S_{k+1} - 2 S_k + S_{k-1} of a source-free step."""
        out = self.zeros(g.nzp, g.nxp)
        self.lib.or_set_operator_only(_I(1))
        try:
            self.update(phys, out, Uc, g, vp, eps, dl, th)
        finally:
            self.lib.or_set_operator_only(_I(0))
        return out

    def adjoint_update(self, phys, Uf, Uc, P, g: Grid, vp, eps=None, dl=None, th=None):
        L = self.lib
        if phys == 1:
            L.or_adjoint_vti(_ptr(Uf), _ptr(Uc), _ptr(P), _ptr(vp), _ptr(eps), _ptr(dl), _I(g.nzp),
                             _I(g.nxp), _D(g.dz), _D(g.dx), _D(g.dt))
        elif phys == 2:
            L.or_adjoint_tti(_ptr(Uf), _ptr(Uc), _ptr(P), _ptr(vp), _ptr(eps), _ptr(dl), _ptr(th),
                             _I(g.nzp), _I(g.nxp), _D(g.dz), _D(g.dx), _D(g.dt))
        else:
            self.update(0, Uf, Uc, g, vp)

    def cerjan_apply(self, U, g: Grid, A):
        self.lib.or_cerjan_apply(_ptr(U), _I(g.nzp), _I(g.nxp), _I(g.N), _ptr(A))

    def grad_vti(self, P, adj, geps, gdel, nz, nx, dx, dz, eps, dl):
        self.lib.or_grad_vti(_ptr(P), _ptr(adj), _ptr(geps), _ptr(gdel), _I(nz), _I(nx), _D(dx), _D(dz),
                             _ptr(eps), _ptr(dl))

    def grad_tti(self, P, adj, geps, gdel, gth, nz, nx, dx, dz, eps, dl, th):
        self.lib.or_grad_tti(_ptr(P), _ptr(adj), _ptr(geps), _ptr(gdel), _ptr(gth), _I(nz), _I(nx),
                             _D(dx), _D(dz), _ptr(eps), _ptr(dl), _ptr(th))

    # whole steps -----------------------------------------------------------
    def make_state(self, g: Grid, abc, vp_exp=None, d0=None, fpico=None):
        st = dict(U0=self.zeros(g.nzp, g.nxp), U1=self.zeros(g.nzp, g.nxp), cp=None)
        if abc == 1:
            prof = self.cpml_profiles(vp_exp, g.N, g.dx, g.dz, g.dt, d0, fpico)
            st["cp"] = prof + [self.zeros(g.nzp, g.nxp) for _ in range(4)]
        return st

    def _cp_ptrs(self, cp):
        if cp is None:
            return None
        arr = (ctypes.c_void_p * 8)(*[c.ctypes.data for c in cp])
        return arr

    def forward_step(self, phys, abc, Uf, Uc, g: Grid, model, A, cp, sz, sx, srcval):
        vp, eps, dl, th = model
        cpp = self._cp_ptrs(cp)
        self.lib.or_forward_step(_I(phys), _I(abc), _ptr(Uf), _ptr(Uc), _ptr(vp), _ptr(eps), _ptr(dl),
                                 _ptr(th), _ptr(A), cpp, _I(g.nzp), _I(g.nxp), _I(g.N), _D(g.dz),
                                 _D(g.dx), _D(g.dt), _I(sz), _I(sx), _D(srcval))

    def forward_shot(self, phys, abc, g: Grid, model, src, sz, sx, rz, rx, A=None, cpml=None,
                     save=False, nsteps=None):
        """Whole forward shot (src/Modeling2D.py:446-460): returns the seismogram
        [nt_data][nrec], the final (current, future) fields, and the padded
        current field after every step when ``save``."""
        vp, eps, dl, th = [self.arr(m) if m is not None else None for m in model]
        nt = g.nt if nsteps is None else nsteps
        st = self.make_state(g, abc, vp, *(cpml or (None, None)))
        if A is None:
            A = self.zeros(max(g.N, 1))
        A = self.arr(A)
        rz = np.ascontiguousarray(rz, dtype=np.int32)
        rx = np.ascontiguousarray(rx, dtype=np.int32)
        nrec = len(rz)
        seis = self.zeros(max(nt - g.itlag, 0), nrec)
        sv = self.zeros(nt, g.nzp, g.nxp) if save else None
        src = self.arr(src)
        self.lib.or_forward_shot(_I(phys), _I(abc), _ptr(st["U0"]), _ptr(st["U1"]), _ptr(vp), _ptr(eps),
                                 _ptr(dl), _ptr(th), _ptr(A), self._cp_ptrs(st["cp"]), _I(g.nzp),
                                 _I(g.nxp), _I(g.N), _D(g.dz), _D(g.dx), _D(g.dt), _I(sz), _I(sx),
                                 _ptr(src), _I(nt), _I(g.itlag), _I(nrec), _ptr(rz), _ptr(rx),
                                 _ptr(seis), _ptr(sv))
        cur, fut = (st["U0"], st["U1"]) if nt % 2 == 0 else (st["U1"], st["U0"])
        return dict(seis=seis, current=cur, future=fut, save=sv, cp=st["cp"])

    def backward_step(self, phys, abc, Uf, Uc, g: Grid, model, A, cp, P, rz, rx, data_row, dxdz):
        """src/Migration.py:264-286: inject data/(dx dz) at receivers (if given),
        then the adjoint update (reference operator choice) and the ABC."""
        vp, eps, dl, th = model
        if data_row is not None:
            rz = np.ascontiguousarray(rz, dtype=np.int32)
            rx = np.ascontiguousarray(rx, dtype=np.int32)
            self.lib.or_inject_rec(_ptr(Uc), _I(g.nxp), _I(len(rz)), _ptr(rz), _ptr(rx),
                                   _ptr(self.arr(data_row)), _D(dxdz))
        if abc == 0:
            if phys == 0:
                self.update(0, Uf, Uc, g, vp)
            else:
                self.adjoint_update(phys, Uf, Uc, P, g, vp, eps, dl, th)
            self.cerjan_apply(Uf, g, A)
            self.cerjan_apply(Uc, g, A)
        else:
            # CPML: the forward operator on the adjoint field (src/Migration.py:270-279)
            self.forward_step(phys, 1, Uf, Uc, g, model, A, cp, -1, -1, 0.0)


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def _bf16_round(x):
    """float32 -> bfloat16 (round to nearest even) -> float32, as the device's
    float -> bfloat16 conversion does for finite values."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def quantize_scaled(v, kind):
    """Values of one step stored in float16 / bfloat16 as v / m with
    m = max |v| (float32 arithmetic: inv = 1/m, q = v * inv), read back as
    q * m; ``kind`` 'f32' returns v unchanged (in its own precision)."""
    if kind == "f32":
        return np.array(v, copy=True)
    v = np.asarray(v, dtype=np.float32)
    m = np.float32(np.max(np.abs(v))) if v.size else np.float32(0.0)
    inv = np.float32(1.0) / m if m >= np.finfo(np.float32).tiny else np.float32(0.0)
    q = v * inv
    q = q.astype(np.float16).astype(np.float32) if kind == "f16" else _bf16_round(q)
    return q * m


def rim_mask(nz, nx):
    """The 4-cell rims just inside the interior edge (src/Migration.py:288-298:
    top / bottom 4 rows, left / right 4 columns of the interior)."""
    m = np.zeros((nz, nx), dtype=bool)
    m[:4, :] = m[nz - 4:, :] = True
    m[:, :4] = m[:, nx - 4:] = True
    return m


def quantize_with_scale(v, m, kind):
    """v stored in float16 / bfloat16 as v * (1/m) for a given scale m (float32
    arithmetic; m <= 0 stores zeros), read back as q * m."""
    v = np.asarray(v, dtype=np.float32)
    m = np.float32(m)
    inv = np.float32(1.0) / m if m >= np.finfo(np.float32).tiny else np.float32(0.0)
    q = v * inv
    q = q.astype(np.float16).astype(np.float32) if kind == "f16" else _bf16_round(q)
    return q * m


class _SaveAll:
    """Source field kept at every step (src/Migration.py:996-1037).

    ``kind`` 'f16' / 'bf16' restates the library's 16-bit scaled snapshots
    (no counterpart in the reference): an image reads S_t stored as S_t / m_t
    with m_t = max |S_t| over the interior; a gradient reads the second time
    difference D_t = S_{t+1} - 2 S_t + S_{t-1} (formed in float32 on the
    interior) stored with the scale of the step before, max |D_{t-1}| (the
    producing step finds the max only as it writes, so D_t's own max comes one
    step late; S_{-1} = 0)."""

    def __init__(self, S, g: Grid, kind="f32"):
        self.S, self.sl, self.kind = S, (slice(g.N, g.N + g.nz), slice(g.N, g.N + g.nx)), kind

    def _d(self, t):
        S, sl = self.S, self.sl
        prev = S[t - 1][sl] if t >= 1 else np.zeros_like(S[t][sl])
        return S[t + 1][sl] - 2.0 * S[t][sl] + prev

    def get(self, t, need_d):
        S, sl = self.S, self.sl
        if self.kind == "f32":
            return S[t], (self._d(t) if need_d else None)
        if need_d:
            scale = np.max(np.abs(self._d(t - 1)))
            return S[t], quantize_with_scale(self._d(t), scale, self.kind)
        St = np.array(S[t], copy=True)
        St[sl] = quantize_scaled(S[t][sl], self.kind)
        return St, None


def _revolve_beta(s, r):
    b = 1.0
    for i in range(1, r + 1):
        b = b * (s + i) / i
    return b


def revolve_split(n, s):
    """Offset of the first checkpoint of an n-step range with s free slots
    (Griewank's treeverse): for the smallest r with beta(s, r) >= n,
    m = min(beta(s, r - 1), n - 1), at least 1; beta(s, r) = C(s + r, s)."""
    r = 1
    while _revolve_beta(s, r) < n:
        r += 1
    return max(1, int(min(_revolve_beta(s, r - 1), float(n - 1))))


class _Revolve:
    """Binomial (revolve) checkpointing of the source field as the library runs
    it (no counterpart in the reference): whole states (both wavefields and,
    with CPML, psi and zeta after the previous step) kept in ``slots`` slots,
    in float32 or ('f16') each field scaled by its max |value| over the padded
    grid and rounded to float16; the forward pass recomputed from the slots in
    treeverse order, each adjoint step's S_t and D_t taken from the one forward
    step run from state t.  The state is carried as the library carries it:
    the current field holds the source sample of its step, and the previous
    field is stored damped once (the Cerjan factors are applied to it again as
    the next step reads it)."""

    def __init__(self, orc, phys, abc, g: Grid, model, A, cpml, src, sz, sx, t_lo, t_hi, slots, kind):
        self.orc, self.phys, self.abc, self.g, self.model, self.A = orc, phys, abc, g, model, A
        self.sl = (slice(g.N, g.N + g.nz), slice(g.N, g.N + g.nx))
        self.src, self.sz, self.sx, self.kind = src, sz, sx, kind
        st = orc.make_state(g, abc, model[0], *(cpml or (None, None)))
        self.cur, self.prv = st["U0"], st["U1"]
        self.cp = st["cp"]
        self._inject(self.cur, 0)
        self.ws = 0
        self.slots = {}
        self.S, self.D = {}, {}
        self.first = True
        self._prev_int = None
        nslot = max(1, min(slots, t_hi - t_lo + 1))
        self._advance(0, t_lo)
        self._store(t_lo)
        self._rev(t_lo, t_hi + 1, nslot - 1)

    def _inject(self, U, k):
        if 0 <= k < self.g.nt:
            self.orc.lib.or_inject(_ptr(U), _I(self.g.nxp), _I(self.sz), _I(self.sx), _D(float(self.src[k])))

    def _fields(self):
        return [self.cur, self.prv] + (list(self.cp[4:]) if self.abc == 1 else [])

    def _step(self, k, save=False):
        o, g = self.orc, self.g
        sl = self.sl
        if save:
            S = np.array(self.cur, copy=True)
            if self.abc == 0:  # the padded snapshot is the current field damped once more
                o.cerjan_apply(S, g, self.A)
            self.S[k] = S
            prev_int = np.array(self.prv[sl], copy=True)
        if self.abc == 0:
            o.cerjan_apply(self.prv, g, self.A)  # the previous field as the step reads it
            vp, eps, dl, th = self.model
            o.update(self.phys, self.prv, self.cur, g, vp, eps, dl, th)
            o.cerjan_apply(self.prv, g, self.A)
        else:
            o.forward_step(self.phys, 1, self.prv, self.cur, g, self.model, self.A, self.cp, -1, -1, 0.0)
        self._inject(self.prv, k + 1)
        if save:  # D_t on the interior (undamped there), in float32 as the step forms it
            self.D[k] = self.prv[sl] - 2.0 * self.cur[sl] + prev_int
        self.cur, self.prv = self.prv, self.cur
        self.ws = k + 1

    def _advance(self, frm, to):
        if self.ws != frm:
            self._load(frm)
        for k in range(frm, to):
            self._step(k)

    def _store(self, k):
        self.slots[k] = [quantize_scaled(f, "f16") if self.kind == "f16" else np.array(f, copy=True)
                         for f in self._fields()]

    def _load(self, k):
        for dst, src in zip(self._fields(), self.slots[k]):
            dst[...] = src
        self.ws = k

    def _leaf(self, t):
        self._advance(t, t)
        self._step(t, save=True)
        if self.first:  # the first descent runs on to the end of the record (illumination)
            self.first = False
            self.ws = self.g.nt

    def _rev(self, lo, hi, s):
        n = hi - lo
        if n == 1:
            return self._leaf(lo)
        if s == 0:
            for t in range(hi - 1, lo - 1, -1):
                self._advance(lo, t)
                self._leaf(t)
            return
        m = lo + revolve_split(n, s)
        self._advance(lo, m)
        self._store(m)
        self._rev(m, hi, s - 1)
        del self.slots[m]
        self._rev(lo, m, s)

    def get(self, t, need_d):
        return self.S[t], (self.D[t] if need_d else None)


class _BoundaryRebuild:
    """Boundary-saving source field: the rims of every S_k (quantised per step
    as ``kind``: 'f32' | 'f16' | 'bf16'), and S rebuilt backwards in time from
    S_nt, S_{nt-1} by S_{k-1} = v^2 dt^2 L(S_k) + 2 S_k - S_{k+1}
    (+ wav[k+1] at a source cell of the inner region) with the undamped update,
    zero on the absorbing layer, the saved rims written back after each step.
    D_k = S_{k-1} - 2 S_k + S_{k+1} is formed as the step goes as the operator
    term v^2 dt^2 L(S_k) (+ wav[k+1] at the source), which the leapfrog makes
    equal to the difference of the three levels without cancelling them; its
    rim cells come from the saved rims.  src/Migration.py:1266-1410 keeps the same rims; its
    time alignment (the reconstructed field paired with the wrong step) is not
    followed."""

    def __init__(self, orc, phys, g: Grid, model, S, U_nt, src, sz, sx, kind):
        self.orc, self.phys, self.g, self.model = orc, phys, g, model
        self.src, self.sz, self.sx = src, sz, sx
        N = g.N
        self.sl = (slice(N, N + g.nz), slice(N, N + g.nx))
        self.mask = rim_mask(g.nz, g.nx)
        self.rims = [quantize_scaled(S[k][self.sl][self.mask], kind) for k in range(g.nt)]
        self.hi = self._interior_only(U_nt)          # S_{lev+1}
        self.lo = self._interior_only(S[g.nt - 1])   # S_lev
        self.lev = g.nt - 1

    def _interior_only(self, U):
        out = np.zeros_like(U)
        out[self.sl] = U[self.sl]
        return out

    def _step(self, need_d=False):
        g, sl, lev = self.g, self.sl, self.lev
        vp, eps, dl, th = self.model
        Uf, Uc = self.hi, self.lo
        d = self.orc.operator(self.phys, Uc, g, vp, eps, dl, th)[sl].copy() if need_d else None
        self.orc.update(self.phys, Uf, Uc, g, vp, eps, dl, th)  # Uf := S_{lev-1} (undamped)
        keep = Uf[sl].copy()
        Uf[...] = 0
        Uf[sl] = keep
        iz, ix = self.sz - g.N, self.sx - g.N
        if lev + 1 < g.nt and 4 <= iz < g.nz - 4 and 4 <= ix < g.nx - 4:
            Uf[self.sz, self.sx] += self.src[lev + 1]
            if need_d:
                d[iz, ix] += self.src[lev + 1]
        Uf[sl][self.mask] = self.rims[lev - 1]
        if need_d:
            d[self.mask] = (self.rims[lev - 1] - 2.0 * self.rims[lev]) + self.rims[lev + 1]
        self.hi, self.lo = Uc, Uf
        self.lev = lev - 1
        return d

    def get(self, t, need_d):
        while self.lev > t:
            self._step()
        if not need_d:
            return self.lo, None
        d = self._step(True)
        return self.hi, d


class _RandomRebuild:
    """Random-boundary source field (src/Migration.py:1413-1561): the forward
    pass ran undamped on the whole grid, so S is rebuilt backwards in time by
    S_{k-1} = v^2 dt^2 L(S_k) + 2 S_k - S_{k+1} (+ wav[k+1] at the source)
    over the whole grid, from S_nt and S_{nt-1}; nothing is stored.  D_k is
    the operator term v^2 dt^2 L(S_k) (+ wav[k+1] at an interior source)."""

    def __init__(self, orc, phys, g: Grid, model, S, U_nt, src, sz, sx):
        self.orc, self.phys, self.g, self.model = orc, phys, g, model
        self.src, self.sz, self.sx = src, sz, sx
        self.sl = (slice(g.N, g.N + g.nz), slice(g.N, g.N + g.nx))
        self.hi = np.array(U_nt, copy=True)
        self.lo = np.array(S[g.nt - 1], copy=True)
        self.lev = g.nt - 1

    def _step(self, need_d=False):
        g, sl, lev = self.g, self.sl, self.lev
        vp, eps, dl, th = self.model
        Uf, Uc = self.hi, self.lo
        d = self.orc.operator(self.phys, Uc, g, vp, eps, dl, th)[sl].copy() if need_d else None
        self.orc.update(self.phys, Uf, Uc, g, vp, eps, dl, th)
        if lev + 1 < g.nt:
            Uf[self.sz, self.sx] += self.src[lev + 1]
            iz, ix = self.sz - g.N, self.sx - g.N
            if need_d and 0 <= iz < g.nz and 0 <= ix < g.nx:
                d[iz, ix] += self.src[lev + 1]
        self.hi, self.lo = Uc, Uf
        self.lev = lev - 1
        return d

    def get(self, t, need_d):
        while self.lev > t:
            self._step()
        if not need_d:
            return self.lo, None
        d = self._step(True)
        return self.hi, d


def rtm_onthefly(orc: Oracle, phys, abc, g: Grid, model, src, shots, rz, rx, data, A=None, cpml=None,
                 fwi=False, multiparameter=False, model_interior=None, reset_cpml=True, t_range="py",
                 normalize=1, storage="all", rim_kind="f32", snap_kind="f32", slots=50):
    """On-the-fly RTM / FWI gradient of src/Migration.py:963-1111 (GPU path):
    per shot, a forward pass storing the current field after every step and
    accumulating ilum += S^2 over the interior, then a backward pass
    t = nt-2..1 with data injection, image += S_t*lambda_t (RTM) or
    grad += lambda_t*(S_{t+1}-2S_t+S_{t-1})/dt^2 (FWI), and for FWI
    multiparameter the VTI/TTI partials.  Finally every output is divided by
    (ilum + 1e-5 max(ilum)).  ``data[s]`` is the (already muted) data or the
    residual for shot s.  With ``reset_cpml`` the CPML memory variables are
    cleared before the backward pass (the reference keeps the forward pass's
    values, src/Migration.py:1000-1039 only resets per shot).

    ``t_range="cpp"`` and ``normalize=2`` give the C++ driver's variant
    (cpp/Migration.cu:383-394, 470-483): the backward pass runs t = nt-1..0
    (imaging only) and the image is divided by ilum where ilum > 1e-12,
    zero elsewhere.

    ``storage="boundary"`` takes the source field from the boundary-saving
    rebuild (src/Migration.py:1266-1410; ``_BoundaryRebuild``) with rims kept
    as ``rim_kind``; ilum still comes from the forward pass.
    ``storage="random"`` runs the forward pass undamped (``model`` being a
    random-boundary model) and takes the source field from the whole-grid
    rebuild (``_RandomRebuild``); the adjoint keeps ``abc``.
    ``snap_kind`` 'f16' / 'bf16' keeps the save-all snapshots in 16 bits as
    the library does (``_SaveAll``); ``storage="revolve"`` recomputes the
    source field from ``slots`` binomial checkpoints kept in ``snap_kind``
    'f32' or 'f16' (``_Revolve``)."""
    if t_range not in ("py", "cpp") or normalize not in (1, 2):
        raise ValueError("t_range 'py'|'cpp', normalize 1|2")
    if t_range == "cpp" and fwi:
        raise ValueError("the C++ backward range applies to imaging only")
    if storage not in ("all", "boundary", "random", "revolve"):
        raise ValueError("storage 'all'|'boundary'|'random'|'revolve'")
    N = g.N
    vp, eps, dl, th = [orc.arr(m) if m is not None else None for m in model]
    image = np.zeros((g.nz, g.nx), dtype=np.float64)
    ilum = np.zeros((g.nz, g.nx), dtype=np.float64)
    ge = np.zeros((g.nz, g.nx), dtype=orc.dtype)
    gd = np.zeros((g.nz, g.nx), dtype=orc.dtype)
    gt = np.zeros((g.nz, g.nx), dtype=orc.dtype)
    if A is None:
        A = orc.zeros(max(N, 1))
    A = orc.arr(A)
    sl = (slice(N, N + g.nz), slice(N, N + g.nx))
    for s, (sz, sx) in enumerate(shots):
        if storage == "random":  # undamped: unit Cerjan factors, no CPML
            fw = orc.forward_shot(phys, 0, g, (vp, eps, dl, th), src, sz, sx, rz, rx, A=orc.arr(np.ones(max(N, 1))),
                                  cpml=None, save=True)
        else:
            fw = orc.forward_shot(phys, abc, g, (vp, eps, dl, th), src, sz, sx, rz, rx, A=A, cpml=cpml, save=True)
        S = fw["save"]
        ilum_p = np.zeros((g.nz, g.nx), dtype=orc.dtype)
        for k in range(g.nt):
            ilum_p += S[k][sl] * S[k][sl]
        t_hi, t_lo = (g.nt - 2, 1) if t_range == "py" else (g.nt - 1, 0)
        if storage == "all":
            prov = _SaveAll(S, g, snap_kind)
        elif storage == "revolve":
            prov = _Revolve(orc, phys, abc, g, (vp, eps, dl, th), A, cpml, orc.arr(src), sz, sx, t_lo, t_hi,
                            slots, snap_kind)
        elif storage == "random":
            prov = _RandomRebuild(orc, phys, g, (vp, eps, dl, th), S, fw["current"], orc.arr(src), sz, sx)
        else:
            prov = _BoundaryRebuild(orc, phys, g, (vp, eps, dl, th), S, fw["current"], orc.arr(src), sz, sx,
                                    rim_kind)
        cp = fw["cp"]
        if storage == "random" and abc == 1:  # the adjoint's own CPML state (the forward had none)
            cp = orc.make_state(g, 1, vp, *cpml)["cp"]
        if cp is not None and reset_cpml:
            for c in cp[4:]:
                c[...] = 0
        Lc = orc.zeros(g.nzp, g.nxp)
        Lf = orc.zeros(g.nzp, g.nxp)
        img_p = np.zeros((g.nz, g.nx), dtype=orc.dtype)
        dxdz = g.dx * g.dz
        t_hi, t_lo = (g.nt - 2, 1) if t_range == "py" else (g.nt - 1, 0)
        for t in range(t_hi, t_lo - 1, -1):
            St, Dt = prov.get(t, fwi)
            row = data[s][t - g.itlag] if t >= g.itlag else None
            orc.backward_step(phys, abc, Lf, Lc, g, (vp, eps, dl, th), A, cp, St, rz, rx, row, dxdz)
            adj = Lc[sl].copy()
            if fwi:
                d2 = Dt / (g.dt * g.dt)
                img_p += (d2 * adj).astype(orc.dtype)
                if multiparameter and phys == 1:
                    orc.grad_vti(np.ascontiguousarray(St[sl]), np.ascontiguousarray(adj), ge, gd, g.nz, g.nx,
                                 g.dx, g.dz, orc.arr(model_interior[1]), orc.arr(model_interior[2]))
                elif multiparameter and phys == 2:
                    orc.grad_tti(np.ascontiguousarray(St[sl]), np.ascontiguousarray(adj), ge, gd, gt, g.nz,
                                 g.nx, g.dx, g.dz, orc.arr(model_interior[1]), orc.arr(model_interior[2]),
                                 orc.arr(model_interior[3]))
            else:
                img_p += (St[sl] * adj).astype(orc.dtype)
            Lc, Lf = Lf, Lc
        image += img_p
        ilum += ilum_p
    if normalize == 1:
        norm = ilum + 1e-5 * np.max(ilum)
    else:
        norm = np.where(ilum > 1e-12, ilum, np.inf)
    out = dict(image=image / norm, raw_image=image, ilum=ilum)
    if multiparameter:
        out.update(g_eps=ge / norm, g_delta=gd / norm, g_theta=gt / norm)
    return out
