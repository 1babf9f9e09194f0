"""The RTM / FWI-gradient driver: class ``migration`` of src/Migration.py with
the same method names and outputs, running on the MI355X through libpam.

``SolveBackwardWaveEquation`` reads the recorded (or residual) gathers with the
reference's file names, smooths the model and mutes the data as the reference
does for RTM (src/Migration.py:963-1030), and writes

* ``migrated_image_{approx}_Nx{nx}_Nz{nz}.bin`` (RTM, imageFolder), or
* ``gradient_{approx}_Nx{nx}_Nz{nz}.bin`` and, for multiparameter FWI,
  ``epsilon_`` / ``delta_`` / ``theta_gradient_...`` (gradientsFolder),

each raw float32 [nz][nx], normalised by ilum + 1e-5 max(ilum)
(src/Migration.py:1086-1110).

Source-field strategies (Parameters.json "migration"): "onthefly" keeps every
step in HBM (float32, or float16 / bfloat16 with "storage"), "checkpoint"
keeps uniform checkpoints every "step" steps in HBM and recomputes segments,
"SB" keeps the 4-cell rims inside the interior edge of every step (float32,
or float16 / bfloat16 with "storage") and rebuilds the source field backwards
in time (src/Migration.py:1266-1410).  "RBC" randomises the absorbing layer of
the velocity model anew for every shot ("dvel", "ratio"), runs the source
undamped through it and rebuilds the source field by reversing those steps
(src/Migration.py:382-458, 1413-1561).
"""
from __future__ import annotations

import time

import numpy as np

from . import _lib, prep

STORAGE = {"float32": _lib.PAM_STORE_ALL_F32, "float16": _lib.PAM_STORE_ALL_F16,
           "bfloat16": _lib.PAM_STORE_ALL_BF16}
SB_STORAGE = {"float32": _lib.PAM_STORE_BOUNDARY_F32, "float16": _lib.PAM_STORE_BOUNDARY_F16,
              "bfloat16": _lib.PAM_STORE_BOUNDARY_BF16}


class migration:
    def __init__(self, wavefield, parameters):
        self.pmt = parameters
        self.wf = wavefield

    def initializeMigrationfields(self):
        p = self.pmt
        self.migrated_image = np.zeros((p.nz, p.nx), dtype=np.float32)
        self.ilum = np.zeros((p.nz, p.nx), dtype=np.float32)
        if p.fwi and p.multiparameter and p.approximation in ("VTI", "TTI"):
            self.epsilon_grad = np.zeros((p.nz, p.nx), dtype=np.float32)
            self.delta_grad = np.zeros((p.nz, p.nx), dtype=np.float32)
            if p.approximation == "TTI":
                self.theta_grad = np.zeros((p.nz, p.nx), dtype=np.float32)
        print(f"info: Wavefields initialized: {p.nx}x{p.nz}x{p.nt}")

    def loadSeismogram(self, shot):
        p = self.pmt
        if p.fwi:
            f = f"{p.seismogramFolder}residual_shot_{shot + 1}_Nt{p.nt_data}_Nrec{p.Nrec}.bin"
        else:
            f = f"{p.seismogramFolder}seismogram_shot_{shot + 1}_Nt{p.nt_data}_Nrec{p.Nrec}_fcut{p.fcut}.bin"
        return np.fromfile(f, dtype=np.float32).reshape(p.nt_data, p.Nrec)

    def _models(self):
        """Migration models: smoothed slowness (and parameters) for RTM
        (src/Migration.py:967-992), the current models for FWI."""
        p = self.pmt
        wf = self.wf
        if p.fwi:
            vp = getattr(self, "vp", wf.vp)
            eps = getattr(self, "epsilon", getattr(wf, "epsilon", None))
            dl = getattr(self, "delta", getattr(wf, "delta", None))
            th = getattr(self, "theta", getattr(wf, "theta", None))
            return vp, eps, dl, th
        from .engine import smooth_model  # the smoothing runs on the GPU

        water = np.abs(wf.vp - np.min(wf.vp)) < 1e-3
        self.vp = smooth_model(wf.vp, p.sigma, water)
        eps = dl = th = None
        if p.approximation in ("VTI", "TTI"):
            eps = smooth_model(wf.epsilon, p.sigma, np.abs(wf.epsilon - np.min(wf.epsilon)) < 1e-3, True)
            dl = smooth_model(wf.delta, p.sigma, np.abs(wf.delta - np.min(wf.delta)) < 1e-3, True)
        if p.approximation == "TTI":
            th = smooth_model(wf.theta, p.sigma, np.abs(wf.theta - np.min(wf.theta)) < 1e-3, True)
        self.epsilon, self.delta, self.theta = eps, dl, th
        return self.vp, eps, dl, th

    def _configure(self, eng):
        p = self.pmt
        if p.migration == "onthefly":
            eng.set_option(_lib.PAM_OPT_STORAGE, STORAGE[getattr(p, "storage", "float32")])
        elif p.migration == "checkpoint":
            eng.set_option(_lib.PAM_OPT_STORAGE, _lib.PAM_STORE_CHECKPOINT)
            eng.set_option(_lib.PAM_OPT_CKPT_STEP, int(p.step) if p.step else 0)
        elif p.migration == "SB":
            eng.set_option(_lib.PAM_OPT_STORAGE, SB_STORAGE[getattr(p, "storage", "float32")])
        elif p.migration == "RBC":
            eng.set_option(_lib.PAM_OPT_STORAGE, _lib.PAM_STORE_RANDOM_BOUNDARY)
        else:
            raise NotImplementedError(f"migration strategy {p.migration!r} is not provided "
                                      "(use 'onthefly', 'checkpoint', 'SB' or 'RBC')")

    def _random_boundary_runs(self, eng, d, models, rng=None):
        """RBC (src/Migration.py:1413-1561): shot by shot, a new random
        absorbing layer (prep.random_boundary) on the expanded velocity, the
        image (or gradient) and illumination summed over the shots, then
        divided by ilum + 1e-5 max(ilum)."""
        import torch

        p = self.pmt
        N = p.N_abc
        vp, eps, dl, th = models
        vp_exp = prep.expand_model(vp, N)
        rest = [prep.expand_model(m, N) if m is not None else None for m in (eps, dl, th)]
        rng = np.random.default_rng() if rng is None else rng
        eng.set_option(_lib.PAM_OPT_NORMALIZE, 0)
        acc = il = None
        for s in range(p.Nshot):
            vr = prep.random_boundary(vp_exp, N, p.dx, p.dz, p.dvel, p.ratio, rng)
            eng.set_model(vr, *rest, layout=_lib.PAM_PADDED)
            if p.fwi:
                out = eng.gradient(d[s:s + 1], shot0=s, nshots=1)
                img, ilum = out["g_m"], out["ilum"]
            else:
                img, ilum = eng.rtm(d[s:s + 1], shot0=s, nshots=1)
            acc = img.clone() if acc is None else acc + img
            il = ilum.clone() if il is None else il + ilum
        eng.set_option(_lib.PAM_OPT_NORMALIZE, 1)
        out = acc / (il + 1e-5 * torch.max(il))
        return out, il

    def data(self):
        """All shots' adjoint sources on the GPU: the recorded gathers muted
        there (RTM, src/utils.py:30-49) or the residuals (FWI),
        [Nshot][nt_data][Nrec]."""
        import torch

        from .engine import mute

        p = self.pmt
        out = np.zeros((p.Nshot, p.nt_data, p.Nrec), dtype=np.float32)
        for s in range(p.Nshot):
            out[s] = self.loadSeismogram(s)
        d = torch.from_numpy(out).to(self.wf._eng.device if getattr(self.wf, "_eng", None) else "cuda:0")
        if not p.fwi:
            mute(d, p.rec_x, p.rec_z, p.shot_x[: p.Nshot], p.shot_z[: p.Nshot], p.dt, p.shift, p.window, p.v0)
        return d

    # -- per-step adjoint (src/Migration.py:237-286) ---------------------------
    def reset_field(self, shot=0, data=None):
        """Clears the adjoint field of ``shot`` and binds its record ``data``
        ([nt_data][Nrec], host or device; the muted gather for RTM, the
        residual for FWI), with zeroed device image / illumination
        accumulators (``image_dev``, ``ilum_dev``) for backward_stepGPU."""
        import torch

        eng = self.wf._eng
        d = self.loadSeismogram(shot) if data is None else data
        eng.adjoint_reset(shot, d)
        self.image_dev = torch.zeros((eng.nz, eng.nx), dtype=torch.float32, device=eng.device)
        self.ilum_dev = torch.zeros_like(self.image_dev)

    def backward_stepGPU(self, k, P=None):
        """One adjoint step k (pam_adjoint_step): data row k injected at the
        receivers, the adjoint update (the linearised VTI/TTI adjoint reads P,
        the padded source field S_k on the device), the absorbing boundary,
        and with P the imaging image += P lambda_k, ilum += P^2."""
        self.wf._eng.adjoint_step(k, P, self.image_dev if P is not None else None,
                                  self.ilum_dev if P is not None else None)

    backward_step = backward_stepGPU

    def SolveBackwardWaveEquation(self, data=None):
        """RTM image (fwi false) or FWI gradient (fwi true) of every shot.
        ``data`` (host or device [Nshot][nt_data][Nrec]) replaces the file
        reads when given."""
        start = time.time()
        p = self.pmt
        vp, eps, dl, th = self._models()
        eng = self.wf.upload_models(vp, eps, dl, th)
        self._configure(eng)
        d = self.data() if data is None else data
        if p.migration == "RBC":
            import torch

            img, ilum = self._random_boundary_runs(eng, torch.as_tensor(d).to(eng.device), (vp, eps, dl, th))
            self.migrated_image = img.cpu().numpy()
            self.ilum = ilum.cpu().numpy()
            f = (f"{p.gradientsFolder}gradient_{p.approximation}_Nx{p.nx}_Nz{p.nz}.bin" if p.fwi else
                 f"{p.imageFolder}migrated_image_{p.approximation}_Nx{p.nx}_Nz{p.nz}.bin")
        elif p.fwi:
            out = eng.gradient(d, multiparameter=bool(p.multiparameter and p.approximation in ("VTI", "TTI")))
            self.migrated_image = out["g_m"].cpu().numpy()
            self.ilum = out["ilum"].cpu().numpy()
            f = f"{p.gradientsFolder}gradient_{p.approximation}_Nx{p.nx}_Nz{p.nz}.bin"
            if out["g_eps"] is not None:
                self.epsilon_grad = out["g_eps"].cpu().numpy()
                self.delta_grad = out["g_delta"].cpu().numpy()
                self.epsilon_grad.tofile(f"{p.gradientsFolder}epsilon_gradient_{p.approximation}_Nx{p.nx}_Nz{p.nz}.bin")
                self.delta_grad.tofile(f"{p.gradientsFolder}delta_gradient_{p.approximation}_Nx{p.nx}_Nz{p.nz}.bin")
            if out["g_theta"] is not None:
                self.theta_grad = out["g_theta"].cpu().numpy()
                self.theta_grad.tofile(f"{p.gradientsFolder}theta_gradient_{p.approximation}_Nx{p.nx}_Nz{p.nz}.bin")
        else:
            img, ilum = eng.rtm(d)
            self.migrated_image = img.cpu().numpy()
            self.ilum = ilum.cpu().numpy()
            f = f"{p.imageFolder}migrated_image_{p.approximation}_Nx{p.nx}_Nz{p.nz}.bin"
        self.outputFile = f
        self.migrated_image.astype(np.float32).tofile(f)
        print(f"info: Final image saved to {f}")
        print(f"info: Migration solved in {time.time() - start:.2f} seconds")
        return self.migrated_image
