"""The forward-modelling driver: class ``wavefield`` of src/Modeling2D.py with
the same method names and outputs, running every shot on the MI355X through
libpam (pseudoacousticmodeling_amd.engine.Engine).

Typical use, as the reference's src/run_modeling.py:

    pmt = parameters("../inputs/Parameters.json")
    wf = wavefield(pmt)
    wf.createSourceWavelet(); wf.initializeWavefields(); wf.loadModels()
    wf.checkDispersionAndStability(); wf.SolveWaveEquation()

Seismograms are written as raw float32 [nt_data][Nrec] files named
``seismogram_shot_{s+1}_Nt{nt_data}_Nrec{Nrec}_fcut{fcut}.bin`` in
seismogramFolder (src/Modeling2D.py:302-305).
"""
from __future__ import annotations

import time

import numpy as np

from . import prep

PHYSICS = {"acoustic": 0, "VTI": 1, "TTI": 2}
ABCS = {"cerjan": 0, "CPML": 1}


class wavefield:
    def __init__(self, parameters):
        self.pmt = parameters
        self._eng = None
        self._eng_key = None

    # -- set-up (src/Modeling2D.py:27-200) -------------------------------------
    def createSourceWavelet(self):
        """Ricker of peak fcut/3 scaled by 1/(dx dz); the Python reference's
        half-derivative filtered form unless Parameters.json sets
        "wavelet": "cpp" (cpp/Modeling.cu:84-100)."""
        p = self.pmt
        if getattr(p, "wavelet", "python") == "cpp":
            self.source = prep.ricker_cpp(p.nt, p.dt, p.fcut, np.float32(p.tlag), p.dx, p.dz)
        else:
            self.source = prep.ricker_py(p.nt, p.dt, p.fcut, p.tlag, p.dx, p.dz)
        print(f"info: Ricker Source wavelet created: {p.nt} samples")

    def ImportModel(self, filename):
        data = np.fromfile(filename, dtype=np.float32)
        if data.size != self.pmt.nz * self.pmt.nx:
            raise ValueError(f"{filename}: {data.size} floats, expected nz*nx = {self.pmt.nz}*{self.pmt.nx}")
        print(f"info: Imported: {filename}")
        return data.reshape(self.pmt.nz, self.pmt.nx)

    def ExpandModel(self, model_data):
        return prep.expand_model(model_data, self.pmt.N_abc)

    def initializeWavefields(self):
        p = self.pmt
        self.vp = np.zeros([p.nz, p.nx], dtype=np.float32)
        self.seismogram = np.zeros([p.nt_data, p.Nrec], dtype=np.float32)
        if p.approximation in ("VTI", "TTI"):
            self.epsilon = np.zeros([p.nz, p.nx], dtype=np.float32)
            self.delta = np.zeros([p.nz, p.nx], dtype=np.float32)
        if p.approximation == "TTI":
            self.theta = np.zeros([p.nz, p.nx], dtype=np.float32)
        print(f"info: Wavefields initialized: {p.nx}x{p.nz}x{p.nt}")

    def loadModels(self):
        """vp, and epsilon / delta (VTI, TTI) and theta in degrees, converted to
        radians (TTI), as src/Modeling2D.py:99-108."""
        p = self.pmt
        self.vp = self.ImportModel(p.vpFile)
        if p.approximation in ("VTI", "TTI"):
            self.epsilon = self.ImportModel(p.epsilonFile)
            self.delta = self.ImportModel(p.deltaFile)
        if p.approximation == "TTI":
            self.theta = np.radians(self.ImportModel(p.thetaFile)).astype(np.float32)
        print(f"info: Models loaded: {p.nx}x{p.nz}")

    def Reflectioncoefficient(self):
        R = 1e-3 ** (self.pmt.N_abc / 10)
        if self.pmt.N_abc >= 200:
            R = 1e-3 ** (150 / 10)
        return R

    def dampening_const(self):
        return prep.dampening_const(self.pmt.N_abc, self.pmt.fcut)

    def createCerjanVector(self):
        return prep.cerjan_vector(self.pmt.N_abc, getattr(self.pmt, "cerjan_factor", 0.0))

    def checkDispersionAndStability(self):
        """Dispersion / stability limits (src/Modeling2D.py:202-256): the
        acoustic rule, or for VTI / TTI the qP phase-velocity scan and the
        8th-order dt limit with the Sd term (prep.stability_anisotropic)."""
        p = self.pmt
        if p.approximation in ("VTI", "TTI"):
            s = prep.stability_anisotropic(self.vp, self.epsilon, self.delta, p.fcut, p.dx, p.dz, p.dt,
                                           theta=self.theta if p.approximation == "TTI" else None)
            print("info: Dispersion and stability check")
            print(f"info: Minimum qP velocity: {s['vqp_min']:.2f} m/s")
            print(f"info: Maximum qP velocity: {s['vqp_max']:.2f} m/s")
            print(f"info: Maximum epsilon: {s['eps_max']:.4f}")
            print(f"info: Maximum delta: {s['delta_max']:.4f}")
        else:
            s = prep.stability_acoustic(self.vp, p.fcut, p.dx, p.dz, p.dt)
            print("info: Dispersion and stability check")
            print(f"info: Minimum velocity: {float(np.min(self.vp)):.2f} m/s")
            print(f"info: Maximum velocity: {float(np.max(self.vp)):.2f} m/s")
        print(f"info: Maximum frequency: {p.fcut:.2f} Hz")
        print(f"info: Critical dx for dispersion: {s['dx_lim']:.2f} m")
        print(f"info: Critical dz for dispersion: {s['dz_lim']:.2f} m")
        print(f"info: Critical dt for stability: {s['dt_lim']:.5f} s")
        if s["ok"]:
            print("info: Dispersion and stability conditions satisfied.")
        else:
            print("WARNING: Dispersion or stability conditions not satisfied.")
        return s

    # -- the GPU engine ------------------------------------------------------
    def engine(self):
        """The libpam context of this survey (one per grid / physics / boundary
        / wavelet); models are (re)uploaded on every call."""
        from .engine import Engine

        p = self.pmt
        key = (p.nz, p.nx, p.N_abc, p.dz, p.dx, p.dt, p.nt, p.itlag, p.approximation, p.ABC, p.Nshot, p.fcut)
        if self._eng is None or self._eng_key != key:
            if self._eng is not None:
                self._eng.close()
            d0, fp = prep.dampening_const(p.N_abc, p.fcut)
            self._eng = Engine(p.nz, p.nx, p.N_abc, p.dz, p.dx, p.dt, p.nt, p.itlag,
                               physics=PHYSICS[p.approximation], abc=ABCS[p.ABC],
                               cerjan_factor=getattr(p, "cerjan_factor", 0.0), cpml_d0=d0, cpml_f_pico=fp,
                               max_shots=max(1, p.Nshot))
            self._eng.set_geometry(p.sz, p.sx, p.rz, p.rx)
            self._eng_key = key
        self._eng.set_wavelet(self.source)
        return self._eng

    def upload_models(self, vp=None, epsilon=None, delta=None, theta=None):
        eng = self.engine()
        p = self.pmt
        vp = self.vp if vp is None else vp
        e = d = t = None
        if p.approximation in ("VTI", "TTI"):
            e = self.epsilon if epsilon is None else epsilon
            d = self.delta if delta is None else delta
        if p.approximation == "TTI":
            t = self.theta if theta is None else theta
        eng.set_model(vp, e, d, t)
        return eng

    def forward_all(self, vp=None, epsilon=None, delta=None, theta=None):
        """Seismograms of every shot on the device, [Nshot][nt_data][Nrec]."""
        eng = self.upload_models(vp, epsilon, delta, theta)
        return eng.forward(0, self.pmt.Nshot)

    # -- per step (src/Modeling2D.py:363-386) ----------------------------------
    def reset_field(self, shot=0):
        self.upload_models()
        self._eng.reset(shot)
        self._k = 0

    def forward_stepGPU(self, k, seis=None):
        """One step of the shot primed by reset_field: injection, update,
        absorbing boundary, and the receiver gather into ``seis`` (a device
        [nt_data][Nrec] tensor) when k >= itlag."""
        self._eng.forward_step(k, seis)

    # -- whole runs (src/Modeling2D.py:423-475) --------------------------------
    def save_seismogram(self, shot):
        p = self.pmt
        self.seismogramFile = f"{p.seismogramFolder}seismogram_shot_{shot + 1}_Nt{p.nt_data}_Nrec{p.Nrec}_fcut{p.fcut}.bin"
        np.ascontiguousarray(self.seismogram, dtype=np.float32).tofile(self.seismogramFile)
        print(f"info: Seismogram saved to {self.seismogramFile}")

    def _snapshots(self, shot):
        """Interior snapshots every `step` steps up to `last_save`
        (src/Modeling2D.py:266-314), through the per-step interface."""
        import torch

        p = self.pmt
        self.reset_field(shot)
        N = p.N_abc
        seis = torch.zeros((p.nt_data, p.Nrec), dtype=torch.float32, device=self._eng.device)
        for k in range(p.nt):
            self._eng.forward_step(k, seis)
            if k <= p.last_save and p.step > 0 and k % p.step == 0:
                # the field of step k as the reference stores it: after the step, before the swap
                cur = self._eng.get_field(1).cpu().numpy()
                snap = cur[N:N + p.nz, N:N + p.nx]
                f = f"{p.snapshotFolder}{p.approximation}forward_shot_{shot + 1}_Nx{p.nx}_Nz{p.nz}_Nt{p.nt}_frame_{k}.bin"
                np.ascontiguousarray(snap, dtype=np.float32).tofile(f)
        return seis

    def SolveWaveEquation(self):
        start = time.time()
        p = self.pmt
        print(f"info: Solving {p.approximation} wave equation")
        seis = self.forward_all().cpu().numpy()
        for shot in range(p.Nshot):
            self.seismogram = seis[shot]
            self.save_seismogram(shot)
            if p.snap:
                self._snapshots(shot)
        print(f"info: {p.Nshot} shots completed in {time.time() - start:.2f} seconds")
        print("info: Wave equation solved")
        return seis

    solveWaveEquationGPU = SolveWaveEquation
    solveWaveEquation = SolveWaveEquation
