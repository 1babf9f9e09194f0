"""The FWI driver: class ``fwi`` of src/Inversion.py with the same method
names, running the forward passes and gradients on the MI355X through libpam.

The parameter is the squared slowness m = 1/vp^2 (src/Inversion.py:15,56,301).
The objective is X = sum_shots 1/2 sum (dobs - dcal)^2, reduced in float64 on
the device (pam_misfit).  The gradient is the migration driver's FWI gradient
(ilum-normalised), zeroed in the water (true vp == 1500, src/Inversion.py:70).
The update is L-BFGS (two-loop recursion, memory 8, src/Inversion.py:88-123,
391-393) with Armijo backtracking (c1 = 1e-4, alpha0 = 0.01 (m_max - m_min),
halving, 10 trials, stop when a trial does not improve on the previous,
src/Inversion.py:125-193).  Every inner product of the optimiser is formed in
float64.

Outputs (src/Inversion.py:353,399-406): per iteration the gradient
``vp_gradient_fwi_iter_{i}_{approx}_Nx{nx}_Nz{nz}_freq{f}.bin`` and the model
``fwi_vp_{approx}_Nx{nx}_Nz{nz}_itr{i}_freq{f}.bin`` (velocity, not m), the
smoothed starting model, and ``history.txt`` (rows X/X0, fmax).
"""
from __future__ import annotations

import os
import time

import numpy as np




def _dot(a, b):
    return float(np.sum(np.asarray(a, dtype=np.float64) * np.asarray(b, dtype=np.float64)))


def two_loop_recursion(g, s_store, y_store):
    """L-BFGS inverse-Hessian product H g (src/Inversion.py:88-123), float64."""
    q = np.asarray(g, dtype=np.float64).copy()
    alpha = np.zeros(len(s_store))
    for i in reversed(range(len(s_store))):
        rho = 1.0 / _dot(s_store[i], y_store[i])
        alpha[i] = rho * _dot(s_store[i], q)
        q = q - alpha[i] * np.asarray(y_store[i], dtype=np.float64)
    gamma = 1.0
    if len(s_store) > 0:
        gamma = _dot(s_store[-1], y_store[-1]) / _dot(y_store[-1], y_store[-1])
    r = gamma * q
    for i in range(len(s_store)):
        rho = 1.0 / _dot(s_store[i], y_store[i])
        beta = rho * _dot(y_store[i], r)
        r = r + np.asarray(s_store[i], dtype=np.float64) * (alpha[i] - beta)
    return r


class fwi:
    def __init__(self, parameters, wavefield, migration, history_file=None):
        self.pmt = parameters
        self.wf = wavefield
        self.mig = migration
        self.history_file = history_file or os.path.join(os.path.dirname(os.path.abspath(parameters.parameters_path)),
                                                         "..", "outputs", "history.txt")

    # -- misfit and gradient -------------------------------------------------
    def loadObsSeismogram(self, shot):
        p = self.pmt
        f = f"{p.seismogramFolder}seismogram_shot_{shot + 1}_Nt{p.nt_data}_Nrec{p.Nrec}_fcut{p.fcut}.bin"
        return np.fromfile(f, dtype=np.float32).reshape(p.nt_data, p.Nrec)

    def save_residual(self, shot, residual):
        p = self.pmt
        f = f"{p.seismogramFolder}residual_shot_{shot + 1}_Nt{p.nt_data}_Nrec{p.Nrec}.bin"
        np.ascontiguousarray(residual, dtype=np.float32).tofile(f)

    def objective_function(self, vp, epsilon, delta, theta, save_residual):
        """X(m) for m = ``vp`` (squared slowness); keeps the residual on the
        device for the gradient and writes the residual files when asked."""
        p = self.pmt
        v = (1.0 / np.sqrt(vp)).astype(np.float32)
        self.vp = v
        dcal = self.wf.forward_all(v, epsilon, delta, theta)
        if getattr(self, "_dobs", None) is None:
            self._dobs = np.stack([self.loadObsSeismogram(s) for s in range(p.Nshot)])
        X, res = self.wf.engine().misfit(self._dobs, dcal)
        self._residual = res
        if save_residual:
            r = res.cpu().numpy()
            for s in range(p.Nshot):
                self.save_residual(s, r[s])
        return X

    def calculate_gradient(self, vp, epsilon, delta, theta):
        """Gradient w.r.t. m at the current residual (src/Inversion.py:55-86)."""
        p = self.pmt
        self.mig.vp = (1.0 / np.sqrt(vp)).astype(np.float32)
        self.mig.epsilon, self.mig.delta, self.mig.theta = epsilon, delta, theta
        p.fwi = True
        self.mig.SolveBackwardWaveEquation(data=self._residual)
        water = np.abs(self.wf.vp - 1500) < 1e-3
        g_vp = self.mig.migrated_image.copy()
        g_vp[water] = 0.0
        g_eps = g_delta = g_theta = None
        if p.multiparameter and p.approximation in ("VTI", "TTI"):
            g_eps = self.mig.epsilon_grad.copy()
            g_delta = self.mig.delta_grad.copy()
            g_eps[water] = 0.0
            g_delta[water] = 0.0
            if p.approximation == "TTI":
                g_theta = self.mig.theta_grad.copy()
                g_theta[water] = 0.0
        return g_vp, g_eps, g_delta, g_theta

    two_loop_recursion = staticmethod(two_loop_recursion)

    def step_length(self, vp, epsilon, delta, theta, p_vp, g_vp, X):
        """Armijo backtracking on m (src/Inversion.py:125-193)."""
        p = self.pmt
        c1 = 1e-4
        m_min = 1.0 / (p.vmax * p.vmax)
        m_max = 1.0 / (p.vmin * p.vmin)
        gTp0 = _dot(g_vp, p_vp)
        alpha = 0.01 * (m_max - m_min)
        X_check = None
        for _ in range(10):
            vp_new = np.clip(vp + alpha * p_vp, m_min, m_max)
            X_new = self.objective_function(vp_new, epsilon, delta, theta, save_residual=False)
            if X_new <= X + c1 * alpha * gTp0:
                return alpha
            if X_check is not None and X_new >= X_check:
                print("info: Armijo condition not reducing objective function. Stopping line search.")
                break
            X_check = X_new
            alpha *= 0.5
        return alpha

    def adjustfmax(self, fmax):
        """Band change: new fcut, tlag, itlag, nt and wavelet (src/Inversion.py:249-258)."""
        p = self.pmt
        p.fcut = fmax
        p.update_sizes()
        self.wf.createSourceWavelet()
        self._dobs = None

    # -- the loop (src/Inversion.py:290-411) -----------------------------------
    def solveFullWaveformInversionMonoparameter(self):
        from .engine import smooth_model  # the smoothing runs on the GPU

        start = time.time()
        p = self.pmt
        mask = np.abs(self.wf.vp - np.min(self.wf.vp)) < 1e-3
        self.vp0 = smooth_model(self.wf.vp, p.sigma, mask)
        self.vp0.astype(np.float32).tofile(f"{p.modelFolder}fwi_vp_smooth_{p.approximation}_Nx{p.nx}_Nz{p.nz}.bin")
        m = (1.0 / (self.vp0.astype(np.float64) ** 2)).astype(np.float32)
        epsilon = delta = theta = None
        if p.approximation in ("VTI", "TTI"):
            epsilon = smooth_model(self.wf.epsilon, p.sigma,
                                   np.abs(self.wf.epsilon - np.min(self.wf.epsilon)) < 1e-3, True)
            delta = smooth_model(self.wf.delta, p.sigma, np.abs(self.wf.delta - np.min(self.wf.delta)) < 1e-3,
                                 True)
        if p.approximation == "TTI":
            theta = smooth_model(self.wf.theta, p.sigma, np.abs(self.wf.theta - np.min(self.wf.theta)) < 1e-3,
                                 True)
        self.history = []
        m_min = 1.0 / (p.vmax * p.vmax)
        m_max = 1.0 / (p.vmin * p.vmin)
        for fmax in p.freqs:
            print(f"info: FWI frequency {fmax} of {p.freqs}")
            self.adjustfmax(fmax)
            s_store, y_store = [], []
            X = self.objective_function(m, epsilon, delta, theta, save_residual=True)
            g, _, _, _ = self.calculate_gradient(m, epsilon, delta, theta)
            X0 = X
            self.history.append([X / X0, fmax])
            for itr in range(p.niter):
                print(f"info: FWI iteration {itr + 1}/{p.niter} for frequency {fmax}")
                g.astype(np.float32).tofile(f"{p.gradientsFolder}vp_gradient_fwi_iter_{itr + 1}_{p.approximation}_"
                                            f"Nx{p.nx}_Nz{p.nz}_freq{fmax}.bin")
                d = -two_loop_recursion(g, s_store, y_store)
                d = (d / np.max(np.abs(d))).astype(np.float32)
                alpha = self.step_length(m, epsilon, delta, theta, d, g, X)
                m_new = np.clip(m + alpha * d, m_min, m_max).astype(np.float32)
                X_new = self.objective_function(m_new, epsilon, delta, theta, save_residual=True)
                g_new, _, _, _ = self.calculate_gradient(m_new, epsilon, delta, theta)
                self.history.append([X_new / X0, fmax])
                s = (m_new - m).astype(np.float64)
                y = (g_new - g).astype(np.float64)
                if _dot(s, y) > 0:
                    s_store.append(s)
                    y_store.append(y)
                if len(s_store) > 8:
                    s_store.pop(0)
                    y_store.pop(0)
                m, g, X = m_new, g_new, X_new
                v_it = (1.0 / np.sqrt(m)).astype(np.float32)
                v_it.tofile(f"{p.estimatedmodelsFolder}fwi_vp_{p.approximation}_Nx{p.nx}_Nz{p.nz}_itr{itr + 1}_freq{fmax}.bin")
        os.makedirs(os.path.dirname(os.path.abspath(self.history_file)), exist_ok=True)
        np.savetxt(self.history_file, np.array(self.history, dtype=np.float32))
        print(f"info: FWI history saved to {self.history_file}")
        print(f"info: FWI finished in {time.time() - start:.2f} s")
        self.m = m
        return m

    def solveFullWaveformInversion(self):
        if self.pmt.multiparameter:
            raise NotImplementedError("the multiparameter FWI schedules are not provided; the multiparameter "
                                      "gradients are (migration.SolveBackwardWaveEquation with multiparameter)")
        return self.solveFullWaveformInversionMonoparameter()
