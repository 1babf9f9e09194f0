"""Thin Python owner of a libpam context.

``Engine`` holds one ``pam_ctx`` and passes torch tensors (device memory and
the current torch stream) through the C ABI.  It is the layer the reference
driver facades (``wavefield``, ``migration``, ``fwi``) and the autograd
function sit on.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib as L


def _f32(t, device):
    if t is None:
        return None
    if not torch.is_tensor(t):
        t = torch.as_tensor(np.ascontiguousarray(t, dtype=np.float32))
    return t.to(device=device, dtype=torch.float32).contiguous()


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


class Engine:
    """One grid, one physics, one absorbing boundary, up to ``max_shots`` shots
    advanced together."""

    def __init__(self, nz, nx, n_abc, dz, dx, dt, nt, itlag, physics="acoustic", abc="cerjan",
                 cerjan_factor=0.0, cpml_d0=0.0, cpml_f_pico=0.0, max_shots=1, device="cuda:0"):
        self.lib = L.load()
        if not torch.cuda.is_available():
            raise RuntimeError("libpam needs a GPU (torch.cuda.is_available() is False)")
        self.device = torch.device(device)
        phys = {"acoustic": L.PAM_ACOUSTIC, "VTI": L.PAM_VTI, "TTI": L.PAM_TTI}.get(physics, physics)
        ab = {"cerjan": L.PAM_CERJAN, "CPML": L.PAM_CPML}.get(abc, abc)
        self.physics, self.abc = phys, ab
        self.grid = L.PamGrid(int(nz), int(nx), int(n_abc), float(dz), float(dx), float(dt), int(nt), int(itlag))
        self.params = L.PamAbcParams(float(cerjan_factor), float(cpml_d0), float(cpml_f_pico))
        self.nz, self.nx, self.N = int(nz), int(nx), int(n_abc)
        self.nzp, self.nxp = self.nz + 2 * self.N, self.nx + 2 * self.N
        self.nt, self.itlag = int(nt), int(itlag)
        self.nt_data = self.nt - self.itlag
        self.max_shots = int(max_shots)
        self.nrec = 0
        self.nshot = 0
        with torch.cuda.device(self.device):
            stream = torch.cuda.current_stream(self.device).cuda_stream
            ctx = ctypes.c_void_p()
            rc = self.lib.pam_create(ctypes.byref(self.grid), phys, ab, ctypes.byref(self.params),
                                     ctypes.c_void_p(stream), self.max_shots, ctypes.byref(ctx))
        self.ctx = ctx
        if rc != L.PAM_OK:
            msg = self.lib.pam_last_error(ctx) if ctx else b"allocation failed"
            if ctx:
                self.lib.pam_destroy(ctx)
                self.ctx = None
            raise L.PamError(f"pam_create: {L.ERRORS.get(rc, rc)}: {msg.decode() if msg else ''}")

    def close(self):
        if getattr(self, "ctx", None):
            self.lib.pam_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ck(self, rc):
        return L.check(self.ctx, rc)

    # -- set-up ----------------------------------------------------------------
    def set_option(self, key, value):
        self._ck(self.lib.pam_set_option(self.ctx, int(key), int(value)))

    def get_option(self, key):
        v = ctypes.c_longlong()
        self._ck(self.lib.pam_get_option(self.ctx, int(key), ctypes.byref(v)))
        return v.value

    def set_model(self, vp, eps=None, delta=None, theta=None, layout=L.PAM_INTERIOR):
        ts = [_f32(x, self.device) for x in (vp, eps, delta, theta)]
        self._model_keep = ts
        self._ck(self.lib.pam_set_model(self.ctx, *[_ptr(t) for t in ts], int(layout)))

    def set_wavelet(self, src):
        t = _f32(src, self.device)
        if t.numel() != self.nt:
            raise ValueError(f"wavelet has {t.numel()} samples, nt = {self.nt}")
        self._ck(self.lib.pam_set_wavelet(self.ctx, _ptr(t)))

    def set_geometry(self, sz, sx, rz, rx):
        a = [np.ascontiguousarray(np.asarray(v).ravel(), dtype=np.int32) for v in (sz, sx, rz, rx)]
        self._geom = a
        p = [x.ctypes.data_as(ctypes.c_void_p) for x in a]
        self.nshot, self.nrec = len(a[0]), len(a[2])
        self._ck(self.lib.pam_set_geometry(self.ctx, p[0], p[1], self.nshot, p[2], p[3], self.nrec))

    # -- per step --------------------------------------------------------------
    def reset(self, shot):
        self._ck(self.lib.pam_reset(self.ctx, int(shot)))

    def forward_step(self, k, seis=None):
        self._ck(self.lib.pam_forward_step(self.ctx, int(k), _ptr(seis)))

    def get_field(self, which=0):
        out = torch.empty((self.nzp, self.nxp), dtype=torch.float32, device=self.device)
        self._ck(self.lib.pam_get_field(self.ctx, int(which), _ptr(out)))
        return out

    def adjoint_reset(self, shot, data):
        """Per-step adjoint of one shot (Migration::backward_step): ``data`` is
        its [nt_data][nrec] record (kept alive until the next reset)."""
        self._adj_data = _f32(data, self.device)
        self._ck(self.lib.pam_adjoint_reset(self.ctx, int(shot), _ptr(self._adj_data)))

    def adjoint_step(self, t, S_t=None, image=None, ilum=None):
        """One adjoint step t: image += S_t lambda_t, ilum += S_t^2 (device
        [nz][nx] tensors, accumulated in place when given)."""
        self._ck(self.lib.pam_adjoint_step(self.ctx, int(t), _ptr(S_t), _ptr(image), _ptr(ilum)))

    # -- whole runs ------------------------------------------------------------
    def forward(self, shot0=0, nshots=None, out=None):
        nshots = self.nshot - shot0 if nshots is None else nshots
        if out is None:
            out = torch.empty((nshots, self.nt_data, self.nrec), dtype=torch.float32, device=self.device)
        self._ck(self.lib.pam_forward(self.ctx, int(shot0), int(nshots), _ptr(out)))
        return out

    def rtm(self, data, shot0=0, nshots=None):
        nshots = data.shape[0] if nshots is None else nshots
        d = _f32(data, self.device)
        image = torch.empty((self.nz, self.nx), dtype=torch.float32, device=self.device)
        ilum = torch.empty_like(image)
        self._ck(self.lib.pam_rtm(self.ctx, int(shot0), int(nshots), _ptr(d), _ptr(image), _ptr(ilum)))
        return image, ilum

    def gradient(self, residual, shot0=0, nshots=None, multiparameter=False):
        nshots = residual.shape[0] if nshots is None else nshots
        d = _f32(residual, self.device)
        g = torch.empty((self.nz, self.nx), dtype=torch.float32, device=self.device)
        ilum = torch.empty_like(g)
        ge = gd = gt = None
        if multiparameter:
            ge = torch.zeros_like(g)
            gd = torch.zeros_like(g)
            if self.physics == L.PAM_TTI:
                gt = torch.zeros_like(g)
        self._ck(self.lib.pam_gradient(self.ctx, int(shot0), int(nshots), _ptr(d), _ptr(g), _ptr(ge), _ptr(gd),
                                       _ptr(gt), _ptr(ilum)))
        return dict(g_m=g, g_eps=ge, g_delta=gd, g_theta=gt, ilum=ilum)

    def misfit(self, dobs, dcal, want_residual=True):
        dobs = _f32(dobs, self.device)
        dcal = _f32(dcal, self.device)
        res = torch.empty_like(dobs) if want_residual else None
        X = ctypes.c_double()
        self._ck(self.lib.pam_misfit(self.ctx, int(dobs.shape[0]), _ptr(dobs), _ptr(dcal), _ptr(res),
                                     ctypes.byref(X)))
        return X.value, res

    def forward_adjoint(self, dseis, shot0=0, nshots=None, layout=L.PAM_INTERIOR, out=None):
        nshots = dseis.shape[0] if nshots is None else nshots
        d = _f32(dseis, self.device)
        if out is None:
            shape = (self.nz, self.nx) if layout == L.PAM_INTERIOR else (self.nzp, self.nxp)
            out = torch.zeros(shape, dtype=torch.float32, device=self.device)
        self._ck(self.lib.pam_forward_adjoint(self.ctx, int(shot0), int(nshots), _ptr(d), _ptr(out), int(layout)))
        return out

    def timing(self):
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_longlong()
        self._ck(self.lib.pam_last_timing(self.ctx, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return dict(loop_ms=a.value, kernel_ms=b.value, kernel_launches=c.value)


class Group:
    """Engines of different grids (the same physics, absorbing boundary and
    nt) advanced together, one launch per time step (pam_group_*)."""

    def __init__(self, engines):
        self.lib = L.load()
        self.engines = list(engines)
        n = len(self.engines)
        arr = (ctypes.c_void_p * n)(*[e.ctx.value for e in self.engines])
        grp = ctypes.c_void_p()
        rc = self.lib.pam_group_create(arr, n, ctypes.byref(grp))
        L.check(self.engines[0].ctx, rc)
        self.grp = grp
        self.loop_ms = 0.0

    def forward(self, nshots, outs=None):
        """Seismograms of shots [0, nshots[i]) of every engine i, as a list of
        device tensors [nshots[i]][nt_data][nrec]."""
        n = len(self.engines)
        if outs is None:
            outs = [torch.empty((int(k), e.nt_data, e.nrec), dtype=torch.float32, device=e.device)
                    for e, k in zip(self.engines, nshots)]
        ns = (ctypes.c_int * n)(*[int(k) for k in nshots])
        ptrs = (ctypes.c_void_p * n)(*[o.data_ptr() for o in outs])
        ms = ctypes.c_double()
        L.check(self.engines[0].ctx, self.lib.pam_group_forward(self.grp, ns, ptrs, ctypes.byref(ms)))
        self.loop_ms = ms.value
        return outs

    def gradient(self, residuals, nshots=None):
        """FWI gradients g_m of every engine for its residual (device
        [nshots[i]][nt_data][nrec]; nshots defaults to each residual's first
        dimension), the members in turn (pam_group_gradient).  Returns (list of
        device [nz][nx] tensors, list of float64 |g_m|^2)."""
        n = len(self.engines)
        if nshots is None:
            nshots = [int(r.shape[0]) for r in residuals]
        res = [_f32(r, e.device) for r, e in zip(residuals, self.engines)]
        outs = [torch.empty((e.nz, e.nx), dtype=torch.float32, device=e.device) for e in self.engines]
        ns = (ctypes.c_int * n)(*[int(k) for k in nshots])
        rp = (ctypes.c_void_p * n)(*[r.data_ptr() for r in res])
        gp = (ctypes.c_void_p * n)(*[o.data_ptr() for o in outs])
        norms = (ctypes.c_double * n)()
        ms = ctypes.c_double()
        L.check(self.engines[0].ctx, self.lib.pam_group_gradient(self.grp, ns, rp, gp, norms, ctypes.byref(ms)))
        self.loop_ms = ms.value
        return outs, [float(v) for v in norms]

    def tile(self):
        return self.lib.pam_group_get_tile(self.grp)

    def close(self):
        if getattr(self, "grp", None):
            self.lib.pam_group_destroy(self.grp)
            self.grp = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def smooth_model(f, sigma, water_mask=None, parameter=False, mode=L.PAM_PREP_PY, device="cuda:0"):
    """Model smoothing before migration on the GPU (pam_smooth_model): the
    Python driver's rule (src/utils.py:76-129) by default, the C++ driver's
    (cpp/Migration.cu:147-236) with ``mode=PAM_PREP_CPP``.  ``f`` is a
    velocity (or, with ``parameter``, an anisotropy parameter) [nz][nx];
    returns the smoothed float32 array on the host."""
    lib = L.load()
    dev = torch.device(device)
    x = _f32(f, dev)
    nz, nx = x.shape
    out = torch.empty_like(x)
    m = None
    if water_mask is not None:
        m = torch.as_tensor(np.ascontiguousarray(water_mask, dtype=np.uint8)).to(dev)
    with torch.cuda.device(dev):
        stream = torch.cuda.current_stream(dev).cuda_stream
        rc = lib.pam_smooth_model(_ptr(x), _ptr(out), _ptr(m), int(nz), int(nx), float(sigma), int(bool(parameter)),
                                  int(mode), ctypes.c_void_p(stream))
    L.check(None, rc)
    return out.cpu().numpy()


def mute(seis, rec_x, rec_z, src_x, src_z, dt, shift, window, v0, mode=L.PAM_PREP_PY):
    """Linear top mute of device gathers ``seis`` [nshots][nt][nrec] in place
    (pam_mute; src/utils.py:30-49 by default, cpp/Migration.cu:86-115 with
    ``mode=PAM_PREP_CPP``); coordinates in metres, one source per shot."""
    lib = L.load()
    ns, nt, nr = seis.shape
    D = ctypes.c_double
    arr = lambda v: (D * len(v))(*[float(t) for t in v])  # noqa: E731
    with torch.cuda.device(seis.device):
        stream = torch.cuda.current_stream(seis.device).cuda_stream
        rc = lib.pam_mute(_ptr(seis), int(ns), int(nt), int(nr), arr(rec_x), arr(rec_z), arr(src_x), arr(src_z),
                          float(dt), float(shift), float(window), float(v0), int(mode), ctypes.c_void_p(stream))
    L.check(None, rc)
    return seis
