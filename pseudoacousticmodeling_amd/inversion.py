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

With ``multiparameter`` (VTI / TTI) the simultaneous schedule runs
(src/Inversion.py:413-655, the reference's dispatch): every parameter
(m, epsilon, delta, and theta for TTI) gets its own L-BFGS direction and
one-parameter Armijo step beta_n, then a joint Armijo step alpha on
m_n + alpha beta_n p_n (src/Inversion.py:195-247).  The hierarchical schedule
(src/Inversion.py:658-940) releases epsilon, delta and theta at 50%, 70% and
85% of the iterations.

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

    # -- multiparameter schedules (src/Inversion.py:195-247, 413-940) ---------
    def _bounds(self):
        p = self.pmt
        return {"vp": (1.0 / (p.vmax * p.vmax), 1.0 / (p.vmin * p.vmin)), "epsilon": (p.epsmin, p.epsmax),
                "delta": (p.deltamin, p.deltamax), "theta": (p.thetamin, p.thetamax)}

    def _objective(self, models, save_residual):
        return self.objective_function(models["vp"], models.get("epsilon"), models.get("delta"),
                                       models.get("theta"), save_residual=save_residual)

    def step_length_parameter(self, name, models, dirs, grads, X):
        """Armijo backtracking on one parameter, the others held
        (src/Inversion.py:125-193): alpha0 = 0.01 (max - min) of its range,
        halved up to 10 times, the trial clipped to the range; stops when a
        trial does not improve on the previous one."""
        c1 = 1e-4
        lo, hi = self._bounds()[name]
        gTp0 = _dot(grads[name], dirs[name])
        alpha = 0.01 * (hi - lo)
        X_check = None
        for _ in range(10):
            trial = dict(models)
            trial[name] = np.clip(models[name] + alpha * dirs[name], lo, hi).astype(np.float32)
            X_new = self._objective(trial, False)
            if X_new <= X + c1 * alpha * gTp0:
                return alpha
            if X_check is not None and X_new >= X_check:
                print("info: Armijo condition not reducing objective function. Stopping line search.")
                break
            X_check = X_new
            alpha *= 0.5
        return alpha

    def step_length_multiparameter(self, models, dirs, grads, betas, X):
        """Armijo backtracking on the joint step m_i + alpha beta_i p_i of every
        parameter (src/Inversion.py:195-247): alpha0 = 1, each trial clipped to
        the parameters' ranges, gT p = sum_i beta_i g_i . p_i (float64)."""
        c1 = 1e-4
        bounds = self._bounds()
        gTp0 = sum(betas[n] * _dot(grads[n], dirs[n]) for n in models)
        alpha = 1.0
        X_check = None
        for _ in range(10):
            trial = {n: np.clip(models[n] + alpha * betas[n] * dirs[n], *bounds[n]).astype(np.float32)
                     for n in models}
            X_new = self._objective(trial, False)
            if X_new <= X + c1 * alpha * gTp0:
                return alpha
            if X_check is not None and X_new >= X_check:
                print("info: Armijo condition not reducing objective function. Stopping line search.")
                break
            X_check = X_new
            alpha *= 0.5
        return alpha

    def _initial_models(self, hierarchical):
        """Starting models (src/Inversion.py:413-465 simultaneous, 666-703
        hierarchical): the simultaneous schedule starts from vp as given and
        smooths epsilon / delta with their non-zero cells kept (the reference
        passes the arrays themselves as masks); the hierarchical one smooths vp
        too and keeps the cells within 1e-3 of each model's minimum.  theta is
        smoothed with the minimum rule in both.  The starting models are
        written as fwi_{vp,epsilon,delta,theta}_smooth_*.bin."""
        from .engine import smooth_model  # the smoothing runs on the GPU

        p = self.pmt
        wf = self.wf
        near_min = lambda f: np.abs(f - np.min(f)) < 1e-3  # noqa: E731
        vp0 = smooth_model(wf.vp, p.sigma, near_min(wf.vp)) if hierarchical else np.asarray(wf.vp, np.float32)
        models = {"vp": (1.0 / (vp0.astype(np.float64) ** 2)).astype(np.float32)}
        out = {"vp": vp0}
        for name, f in (("epsilon", wf.epsilon), ("delta", wf.delta)):
            mask = near_min(f) if hierarchical else (f != 0)
            models[name] = out[name] = smooth_model(f, p.sigma, mask, True)
        if p.approximation == "TTI":
            models["theta"] = out["theta"] = smooth_model(wf.theta, p.sigma, near_min(wf.theta), True)
        for name, f in out.items():
            f.astype(np.float32).tofile(f"{p.modelFolder}fwi_{name}_smooth_{p.approximation}_Nx{p.nx}_Nz{p.nz}.bin")
        return models

    def _solve_multiparameter(self, release, hierarchical):
        """One loop for both schedules: parameter n starts updating once the
        iteration's progress (itr + 1) / niter reaches release[n] (vp at 0).
        Each active parameter gets its own L-BFGS direction (memory 8) scaled
        to max |p| = 1 and its own step beta_n from a one-parameter Armijo
        search; the joint step alpha comes from the multiparameter search (1
        while only vp is active).  Per iteration the active parameters'
        gradients and models are written (vp as a velocity), and history.txt
        holds X/X0 per iteration."""
        start = time.time()
        p = self.pmt
        if p.approximation not in ("VTI", "TTI"):
            raise ValueError("multiparameter FWI inverts epsilon / delta (and theta): it needs VTI or TTI")
        models = self._initial_models(hierarchical)
        names = list(models)
        bounds = self._bounds()
        self.history = []

        def grads_of(m):
            g = self.calculate_gradient(m["vp"], m["epsilon"], m["delta"], m.get("theta"))
            return dict(zip(["vp", "epsilon", "delta", "theta"], g))

        for fmax in p.freqs:
            print(f"info: FWI frequency {fmax} of {p.freqs}")
            self.adjustfmax(fmax)
            mem = {n: ([], []) for n in names}
            X = self._objective(models, True)
            grads = grads_of(models)
            X0 = X
            self.history.append([X / X0, fmax])
            for itr in range(p.niter):
                progress = (itr + 1) / p.niter
                active = [n for n in names if progress >= release[n]]
                print(f"info: FWI iteration {itr + 1}/{p.niter} for frequency {fmax}, updating {', '.join(active)}")
                for n in active:
                    tag = "vp" if n == "vp" else n
                    grads[n].astype(np.float32).tofile(f"{p.gradientsFolder}{tag}_gradient_fwi_iter_{itr + 1}_"
                                                       f"{p.approximation}_Nx{p.nx}_Nz{p.nz}_freq{fmax}.bin")
                dirs, betas = {}, {}
                for n in names:
                    if n in active:
                        d = -two_loop_recursion(grads[n], *mem[n])
                        dirs[n] = (d / np.max(np.abs(d))).astype(np.float32)
                    else:
                        dirs[n] = np.zeros_like(models[n])
                for n in names:
                    betas[n] = self.step_length_parameter(n, models, dirs, grads, X) if n in active else 0.0
                if active == ["vp"]:
                    alpha = 1.0
                else:
                    alpha = self.step_length_multiparameter(models, dirs, grads, betas, X)
                new = {n: np.clip(models[n] + alpha * betas[n] * dirs[n], *bounds[n]).astype(np.float32)
                       for n in names}
                X_new = self._objective(new, True)
                grads_new = grads_of(new)
                self.history.append([X_new / X0, fmax])
                for n in active:
                    s_ = (new[n] - models[n]).astype(np.float64)
                    y_ = (grads_new[n] - grads[n]).astype(np.float64)
                    if _dot(s_, y_) > 0:
                        mem[n][0].append(s_)
                        mem[n][1].append(y_)
                    if len(mem[n][0]) > 8:
                        mem[n][0].pop(0)
                        mem[n][1].pop(0)
                for n in active:
                    models[n] = new[n]
                    grads[n] = grads_new[n]
                X = X_new
                for n in active:
                    out = (1.0 / np.sqrt(models[n])).astype(np.float32) if n == "vp" else models[n]
                    out.tofile(f"{p.estimatedmodelsFolder}fwi_{n}_{p.approximation}_Nx{p.nx}_Nz{p.nz}_"
                               f"itr{itr + 1}_freq{fmax}.bin")
        os.makedirs(os.path.dirname(os.path.abspath(self.history_file)), exist_ok=True)
        np.savetxt(self.history_file, np.array(self.history, dtype=np.float32))
        print(f"info: FWI history saved to {self.history_file}")
        print(f"info: FWI finished in {time.time() - start:.2f} s")
        self.models = models
        return models

    def solveFullWaveformInversionMultiparameterSimultaneos(self):
        """Every parameter updated from the first iteration
        (src/Inversion.py:413-655)."""
        return self._solve_multiparameter({"vp": 0.0, "epsilon": 0.0, "delta": 0.0, "theta": 0.0}, False)

    def solveFullWaveformInversionMultiparameterHierarchical(self):
        """vp first; epsilon from half the iterations, delta from 70%, theta
        from 85% (src/Inversion.py:658-940)."""
        return self._solve_multiparameter({"vp": 0.0, "epsilon": 0.50, "delta": 0.70, "theta": 0.85}, True)

    def solveFullWaveformInversion(self):
        """src/Inversion.py:942-946: the simultaneous schedule when multiparameter."""
        if self.pmt.multiparameter:
            return self.solveFullWaveformInversionMultiparameterSimultaneos()
        return self.solveFullWaveformInversionMonoparameter()
