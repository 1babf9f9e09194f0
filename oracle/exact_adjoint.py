"""Exact discrete adjoint of the acoustic forward map -- TEST INFRASTRUCTURE ONLY.

Only ``tests/`` import this module, as the checker of the library's
``pam_forward_adjoint``.  float64 numpy.

The forward map is the one the library and ``pam_oracle.c`` implement for
acoustic physics (src/Modeling2D.py:363-386, src/utils.py:396-410 and
242-377, src/CudaKernels.py:268-305 and 454-721): with C_k the current field
after the injection of src[k] and before damping,

  Cerjan  C_{k+1} = D (V Q_k + 2 C_k - D C_{k-1}) + e src[k+1],
          Q_k = Dxx C_k + Dzz C_k,                  seis_k = (D C_k)[rec]
  CPML    psi_x,k = Sx (a_x psi_x,k-1 + b_x Dx C_k)
          zeta_x,k = Sx (a_x zeta_x,k-1 + b_x (Dxx C_k + Dx psi_x,k))   (same in z)
          Q_k = Dxx C_k + Dzz C_k + Sx (Dx psi_x,k + zeta_x,k) + Sz (Dz psi_z,k + zeta_z,k)
          C_{k+1} = V Q_k + 2 C_k - C_{k-1} + e src[k+1],   seis_k = C_k[rec]

on the live cells (4 <= i < n-4 in both directions; the rim stays zero),
V = vp^2 dt^2, D the Cerjan factors, Sx / Sz the live CPML strips and
a, b the CPML profiles, which depend on vp (src/CudaKernels.py:33-90).

Three functions:

* ``forward``  -- the map vp (padded) -> seismogram, optionally keeping the
  per-step state the adjoint needs;
* ``tangent``  -- its directional derivative in vp (the Born operator),
  including the profiles' dependence on vp;
* ``adjoint``  -- the transpose of ``tangent`` applied to a seismogram
  residual r: dJ/dvp (padded) for J with dJ/dseis = r.

``expand_T`` is the transpose of the edge-replicating expansion
(src/Modeling2D.py:38-54), taking a padded gradient to the interior model.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

KC = np.array([-1435.0 / 504.0, 8.0 / 5.0, -1.0 / 5.0, 8.0 / 315.0, -1.0 / 560.0])
KA = np.array([4.0 / 5.0, -1.0 / 5.0, 4.0 / 105.0, -1.0 / 280.0])


def _d2(U, axis):
    """Undivided 8th-order second difference along ``axis`` on the live band
    (4 .. n-5 along that axis); zero elsewhere."""
    U = np.moveaxis(U, axis, 0)
    n = U.shape[0]
    out = np.zeros_like(U)
    out[4:n - 4] = KC[0] * U[4:n - 4]
    for m in range(1, 5):
        out[4:n - 4] += KC[m] * (U[4 + m:n - 4 + m] + U[4 - m:n - 4 - m])
    return np.moveaxis(out, 0, axis)


def _d1(U, axis):
    """Undivided 8th-order first difference (antisymmetric) on the live band."""
    U = np.moveaxis(U, axis, 0)
    n = U.shape[0]
    out = np.zeros_like(U)
    for m in range(1, 5):
        out[4:n - 4] += KA[m - 1] * (U[4 + m:n - 4 + m] - U[4 - m:n - 4 - m])
    return np.moveaxis(out, 0, axis)


def _d1T(U, axis):
    """Transpose of ``_d1``: -d1 of the live-band-restricted input, evaluated
    everywhere it reaches."""
    U = np.moveaxis(U, axis, 0)
    n = U.shape[0]
    Ul = np.zeros_like(U)
    Ul[4:n - 4] = U[4:n - 4]
    out = np.zeros_like(U)
    for m in range(1, 5):
        # out[i] = sum_m KA[m] (Ul[i-m] - Ul[i+m])
        out[m:] += KA[m - 1] * Ul[:n - m]
        out[:n - m] -= KA[m - 1] * Ul[m:]
    return np.moveaxis(out, 0, axis)


def _d2T(U, axis):
    """Transpose of ``_d2`` (symmetric stencil on the live-band-restricted input)."""
    U = np.moveaxis(U, axis, 0)
    n = U.shape[0]
    Ul = np.zeros_like(U)
    Ul[4:n - 4] = U[4:n - 4]
    out = KC[0] * Ul
    for m in range(1, 5):
        out[m:] += KC[m] * Ul[:n - m]
        out[:n - m] += KC[m] * Ul[m:]
    return np.moveaxis(out, 0, axis)


@dataclass
class Setup:
    nzp: int
    nxp: int
    N: int
    dz: float
    dx: float
    dt: float
    nt: int
    itlag: int
    abc: int  # 0 Cerjan, 1 CPML
    src: np.ndarray  # [nt]
    sz: int
    sx: int
    rz: np.ndarray
    rx: np.ndarray
    A: np.ndarray = None  # Cerjan taper [N]
    d0: float = 0.0
    fpico: float = 0.0

    def live(self):
        M = np.zeros((self.nzp, self.nxp))
        M[4:self.nzp - 4, 4:self.nxp - 4] = 1.0
        return M

    def damping(self):
        """D = product of the Cerjan factors of the cell's row and column
        (src/CudaKernels.py:3-31)."""
        N = self.N

        def f(n):
            v = np.ones(n)
            if self.abc == 0:
                for i in range(n):
                    if i < N:
                        v[i] *= self.A[i]
                    if i >= n - N:
                        v[i] *= self.A[n - 1 - i]
            return v

        return f(self.nzp)[:, None] * f(self.nxp)[None, :]

    def strips(self):
        N = self.N
        M = self.live()
        ix = np.arange(self.nxp)
        iz = np.arange(self.nzp)
        sx = ((ix < N) | (ix >= self.nxp - N)).astype(float)
        sz = ((iz < N) | (iz >= self.nzp - N)).astype(float)
        return M * sx[None, :], M * sz[:, None]

    def _prof(self, n, h, vp, along_x):
        """(a, b, da/dvp, db/dvp) of the CPML profile along one axis
        (src/CudaKernels.py:33-90) for a padded vp."""
        N = self.N
        idx = np.arange(n, dtype=np.float64)
        pos = np.zeros(n)
        inside = np.zeros(n, dtype=bool)
        lo = idx < N
        hi = idx >= n - N
        pos[lo] = ((N - idx[lo] - 1.0) * h) / (N * h)
        pos[hi] = ((idx[hi] - n + N) * h) / (N * h)
        inside[lo | hi] = True
        kap = np.where(inside, self.d0 / (2.0 * N * h) * pos * pos, 0.0)
        alp = np.where(inside, np.pi * self.fpico * (1.0 - pos * pos), 0.0)
        kap = kap[None, :] if along_x else kap[:, None]
        alp = alp[None, :] if along_x else alp[:, None]
        d = kap * vp
        s = d + alp
        a = np.exp(-s * self.dt)
        ok = np.abs(s) > 1e-10
        ss = np.where(ok, s, 1.0)
        b = np.where(ok, d / ss * (a - 1.0), 0.0)
        da = -kap * self.dt * a
        db = np.where(ok, kap * alp / (ss * ss) * (a - 1.0) + d / ss * da, 0.0)
        return a, b, da, db

    def profiles(self, vp):
        ax, bx, dax, dbx = self._prof(self.nxp, self.dx, vp, True)
        az, bz, daz, dbz = self._prof(self.nzp, self.dz, vp, False)
        return dict(ax=ax, bx=bx, az=az, bz=bz, dax=dax, dbx=dbx, daz=daz, dbz=dbz)


def forward(st: Setup, vp, keep=False, dvp=None):
    """Seismogram [nt-itlag][nrec] of the padded model ``vp``.  With ``keep``,
    also the per-step states (C_k, and psi_k / zeta_k for CPML).  With ``dvp``,
    also the tangent seismogram (directional derivative along dvp)."""
    vp = np.asarray(vp, dtype=np.float64)
    M = st.live()
    V = vp * vp * st.dt * st.dt
    D = st.damping()
    Sx, Sz = st.strips()
    pr = st.profiles(vp) if st.abc == 1 else None
    idx2, idz2 = 1.0 / (st.dx * st.dx), 1.0 / (st.dz * st.dz)
    idx, idz = 1.0 / st.dx, 1.0 / st.dz
    shape = (st.nzp, st.nxp)
    C = np.zeros(shape)
    Cp = np.zeros(shape)
    C[st.sz, st.sx] += st.src[0]
    px = np.zeros(shape); pz = np.zeros(shape); zx = np.zeros(shape); zz = np.zeros(shape)
    tan = dvp is not None
    if tan:
        dvp = np.asarray(dvp, dtype=np.float64)
        dV = 2.0 * vp * dvp * st.dt * st.dt
        dC = np.zeros(shape); dCp = np.zeros(shape)
        dpx = np.zeros(shape); dpz = np.zeros(shape); dzx = np.zeros(shape); dzz = np.zeros(shape)
        dseis = np.zeros((st.nt - st.itlag, len(st.rz)))
    seis = np.zeros((st.nt - st.itlag, len(st.rz)))
    states = []
    for k in range(st.nt):
        lap = _d2(C, 1) * idx2 + _d2(C, 0) * idz2
        if tan:
            dlap = _d2(dC, 1) * idx2 + _d2(dC, 0) * idz2
        if st.abc == 1:
            ax, bx, az, bz = pr["ax"], pr["bx"], pr["az"], pr["bz"]
            gx = _d1(C, 1) * idx
            gz = _d1(C, 0) * idz
            px_n = Sx * (ax * px + bx * gx)
            pz_n = Sz * (az * pz + bz * gz)
            ex = _d1(px_n, 1) * idx
            ez = _d1(pz_n, 0) * idz
            zx_n = Sx * (ax * zx + bx * (_d2(C, 1) * idx2 + ex))
            zz_n = Sz * (az * zz + bz * (_d2(C, 0) * idz2 + ez))
            Q = M * (lap + Sx * (ex + zx_n) + Sz * (ez + zz_n))
            if tan:
                dax, dbx, daz, dbz = pr["dax"] * dvp, pr["dbx"] * dvp, pr["daz"] * dvp, pr["dbz"] * dvp
                dgx = _d1(dC, 1) * idx
                dgz = _d1(dC, 0) * idz
                dpx_n = Sx * (dax * px + ax * dpx + dbx * gx + bx * dgx)
                dpz_n = Sz * (daz * pz + az * dpz + dbz * gz + bz * dgz)
                dex = _d1(dpx_n, 1) * idx
                dez = _d1(dpz_n, 0) * idz
                dzx_n = Sx * (dax * zx + ax * dzx + dbx * (_d2(C, 1) * idx2 + ex) + bx * (_d2(dC, 1) * idx2 + dex))
                dzz_n = Sz * (daz * zz + az * dzz + dbz * (_d2(C, 0) * idz2 + ez) + bz * (_d2(dC, 0) * idz2 + dez))
                dQ = M * (dlap + Sx * (dex + dzx_n) + Sz * (dez + dzz_n))
            if keep:
                states.append(dict(C=C.copy(), px=px_n, pz=pz_n, zx=zx_n, zz=zz_n))
            F = M * (V * Q + 2.0 * C - Cp)
            if tan:
                dF = M * (dV * Q + V * dQ + 2.0 * dC - dCp)
            obs = C
            dobs = dC if tan else None
            px, pz, zx, zz = px_n, pz_n, zx_n, zz_n
            if tan:
                dpx, dpz, dzx, dzz = dpx_n, dpz_n, dzx_n, dzz_n
            Cn = F
            if tan:
                dCn = dF
        else:
            Q = M * lap
            if keep:
                states.append(dict(C=C.copy()))
            F = M * (V * Q + 2.0 * C - D * Cp)
            obs = D * C
            Cn = D * F
            if tan:
                dQ = M * dlap
                dF = M * (dV * Q + V * dQ + 2.0 * dC - D * dCp)
                dobs = D * dC
                dCn = D * dF
        if k >= st.itlag:
            seis[k - st.itlag] = obs[st.rz, st.rx]
            if tan:
                dseis[k - st.itlag] = dobs[st.rz, st.rx]
        if k + 1 < st.nt:
            Cn[st.sz, st.sx] += st.src[k + 1]
        Cp, C = C, Cn
        if tan:
            dCp, dC = dC, dCn
    out = dict(seis=seis)
    if keep:
        out["states"] = states
    if tan:
        out["dseis"] = dseis
    return out


def tangent(st: Setup, vp, dvp):
    return forward(st, vp, dvp=dvp)["dseis"]


def adjoint(st: Setup, vp, r, states=None, prev_states=None):
    """dJ/dvp on the padded grid for dJ/dseis = r ([nt-itlag][nrec]), by the
    reverse sweep of the forward map (k = nt-1 .. 0).  ``states[k]`` is the
    state of step k; the psi / zeta of step k - 1 it also reads come from
    ``prev_states[k]`` when given (a recomputed sweep may hold two versions of
    one step), else from ``states[k - 1]``."""
    vp = np.asarray(vp, dtype=np.float64)
    if states is None:
        states = forward(st, vp, keep=True)["states"]
    M = st.live()
    V = vp * vp * st.dt * st.dt
    D = st.damping()
    Sx, Sz = st.strips()
    idx2, idz2 = 1.0 / (st.dx * st.dx), 1.0 / (st.dz * st.dz)
    idx, idz = 1.0 / st.dx, 1.0 / st.dz
    shape = (st.nzp, st.nxp)
    X = np.zeros(shape)   # lambda_{k+1}
    Y = np.zeros(shape)   # lambda_{k+2}
    gV = np.zeros(shape)
    cp = st.abc == 1
    if cp:
        pr = st.profiles(vp)
        ax, bx, az, bz = pr["ax"], pr["bx"], pr["az"], pr["bz"]
        mux = np.zeros(shape); muz = np.zeros(shape); nux = np.zeros(shape); nuz = np.zeros(shape)
        gax = np.zeros(shape); gbx = np.zeros(shape); gaz = np.zeros(shape); gbz = np.zeros(shape)
    rz, rx = np.asarray(st.rz), np.asarray(st.rx)
    for k in range(st.nt - 1, -1, -1):
        S = states[k]
        C = S["C"]
        if cp:
            phi = M * X
            aQ = V * phi
            nux_k = Sx * (aQ + ax * nux)
            nuz_k = Sz * (aQ + az * nuz)
            mux_k = Sx * (ax * mux + _d1T(Sx * aQ + bx * nux_k, 1) * idx)
            muz_k = Sz * (az * muz + _d1T(Sz * aQ + bz * nuz_k, 0) * idz)
            lam = (_d2T(aQ + bx * nux_k, 1) * idx2 + _d2T(aQ + bz * nuz_k, 0) * idz2 + 2.0 * phi - M * Y
                   + _d1T(bx * mux_k, 1) * idx + _d1T(bz * muz_k, 0) * idz)
            lam = M * lam
            Q = M * (_d2(C, 1) * idx2 + _d2(C, 0) * idz2 + Sx * (_d1(S["px"], 1) * idx + S["zx"])
                     + Sz * (_d1(S["pz"], 0) * idz + S["zz"]))
            gV += phi * Q
            prev = (prev_states[k] if prev_states is not None else states[k - 1]) if k > 0 else None
            pxm = prev["px"] if prev else 0.0
            pzm = prev["pz"] if prev else 0.0
            zxm = prev["zx"] if prev else 0.0
            zzm = prev["zz"] if prev else 0.0
            gax += mux_k * pxm + nux_k * zxm
            gaz += muz_k * pzm + nuz_k * zzm
            gbx += mux_k * (_d1(C, 1) * idx) + nux_k * (_d2(C, 1) * idx2 + _d1(S["px"], 1) * idx)
            gbz += muz_k * (_d1(C, 0) * idz) + nuz_k * (_d2(C, 0) * idz2 + _d1(S["pz"], 0) * idz)
            mux, muz, nux, nuz = mux_k, muz_k, nux_k, nuz_k
            w = 1.0
        else:
            phi = M * D * X
            lam = M * (_d2T(V * phi, 1) * idx2 + _d2T(V * phi, 0) * idz2 + 2.0 * phi - D * (M * D * Y))
            gV += phi * (M * (_d2(C, 1) * idx2 + _d2(C, 0) * idz2))
            w = D
        if k >= st.itlag:
            wr = w[rz, rx] if cp is False else 1.0
            np.add.at(lam, (rz, rx), wr * r[k - st.itlag])
        Y, X = X, lam
    g = gV * 2.0 * vp * st.dt * st.dt
    if cp:
        g = g + gax * pr["dax"] + gbx * pr["dbx"] + gaz * pr["daz"] + gbz * pr["dbz"]
    return g


def _quantize(v, kind):
    """float16 storage of one field as the library keeps a revolve slot (the
    quantiser of oracle.quantize_scaled: v / max|v| in float32, rounded to
    float16, read back times max|v|); 'f32' keeps v."""
    if kind == "f32":
        return np.array(v, dtype=np.float64, copy=True)
    v = np.asarray(v, dtype=np.float32)
    m = np.float32(np.max(np.abs(v))) if v.size else np.float32(0.0)
    inv = np.float32(1.0) / m if m >= np.finfo(np.float32).tiny else np.float32(0.0)
    q = (v * inv).astype(np.float16).astype(np.float32)
    return (q * m).astype(np.float64)


def _revolve_beta(s, r):
    b = 1.0
    for i in range(1, r + 1):
        b = b * (s + i) / i
    return b


def revolve_split(n, s):
    """Offset of the first checkpoint of an n-step range with s free slots
    (Griewank's treeverse, as the library splits): for the smallest r with
    beta(s, r) >= n, min(beta(s, r - 1), n - 1), at least 1."""
    r = 1
    while _revolve_beta(s, r) < n:
        r += 1
    return max(1, int(min(_revolve_beta(s, r - 1), float(n - 1))))


def revolve_states(st: Setup, vp, slots, kind="f16"):
    """The states the library's reverse sweep reads with binomial (revolve)
    checkpointing of the whole forward state in ``slots`` slots stored as
    ``kind`` (no counterpart in the reference): a slot keeps the state before
    a step -- C_k, C_{k-1} and, for CPML, the previous step's psi and zeta --
    each field quantised on its own; the steps are reversed from nt - 1 down
    to 0 in treeverse order, reverse step t reading the state of step t and
    the psi / zeta of step t - 1, both from the one forward recomputation that
    reaches t.  Returns (states, prev_states) for ``adjoint``."""
    vp = np.asarray(vp, dtype=np.float64)
    M = st.live()
    V = vp * vp * st.dt * st.dt
    D = st.damping()
    Sx, Sz = st.strips()
    pr = st.profiles(vp) if st.abc == 1 else None
    idx2, idz2 = 1.0 / (st.dx * st.dx), 1.0 / (st.dz * st.dz)
    idx, idz = 1.0 / st.dx, 1.0 / st.dz
    shape = (st.nzp, st.nxp)
    cp = st.abc == 1
    names = ("C", "Cp", "px", "pz", "zx", "zz") if cp else ("C", "Cp")

    def step(w, k):
        """one forward step of the working state w (in place); returns the
        state of step k (C_k and the psi / zeta step k writes)"""
        C, Cp = w["C"], w["Cp"]
        lap = _d2(C, 1) * idx2 + _d2(C, 0) * idz2
        if cp:
            ax, bx, az, bz = pr["ax"], pr["bx"], pr["az"], pr["bz"]
            px_n = Sx * (ax * w["px"] + bx * (_d1(C, 1) * idx))
            pz_n = Sz * (az * w["pz"] + bz * (_d1(C, 0) * idz))
            ex = _d1(px_n, 1) * idx
            ez = _d1(pz_n, 0) * idz
            zx_n = Sx * (ax * w["zx"] + bx * (_d2(C, 1) * idx2 + ex))
            zz_n = Sz * (az * w["zz"] + bz * (_d2(C, 0) * idz2 + ez))
            Q = M * (lap + Sx * (ex + zx_n) + Sz * (ez + zz_n))
            snap = dict(C=C.copy(), px=px_n, pz=pz_n, zx=zx_n, zz=zz_n)
            Cn = M * (V * Q + 2.0 * C - Cp)
            w["px"], w["pz"], w["zx"], w["zz"] = px_n, pz_n, zx_n, zz_n
        else:
            Q = M * lap
            snap = dict(C=C.copy())
            Cn = D * (M * (V * Q + 2.0 * C - D * Cp))
        if k + 1 < st.nt:
            Cn[st.sz, st.sx] += st.src[k + 1]
        w["Cp"], w["C"] = C, Cn
        return snap

    w = {n: np.zeros(shape) for n in names}
    w["C"][st.sz, st.sx] += st.src[0]
    pos = [0]  # step whose state the working set holds
    store, states, prevs = {}, [None] * st.nt, [None] * st.nt

    def save(k):
        store[k] = {n: _quantize(w[n], kind) for n in names}

    def advance(frm, to):
        if pos[0] != frm:
            for n in names:
                w[n] = store[frm][n].copy()
            pos[0] = frm
        for k in range(frm, to):
            step(w, k)
            pos[0] = k + 1

    def leaf(t):
        advance(t, t)
        if t >= 1 and cp:
            prevs[t] = {n: w[n].copy() for n in ("px", "pz", "zx", "zz")}
        states[t] = step(w, t)
        pos[0] = t + 1

    def rev(lo, hi, s):
        n = hi - lo
        if n == 1:
            return leaf(lo)
        if s == 0:
            for t in range(hi - 1, lo - 1, -1):
                advance(lo, t)
                leaf(t)
            return
        m = lo + revolve_split(n, s)
        advance(lo, m)
        save(m)
        rev(m, hi, s - 1)
        del store[m]
        rev(lo, m, s)

    save(0)
    rev(0, st.nt, max(1, min(slots, st.nt)) - 1)
    return states, (prevs if cp else None)


def expand_T(gpad, N):
    """Transpose of the edge-replicating pad by N: each interior cell collects
    the padded cells that copy it."""
    g = np.asarray(gpad, dtype=np.float64)
    nzp, nxp = g.shape
    nz, nx = nzp - 2 * N, nxp - 2 * N
    rows = np.clip(np.arange(nzp) - N, 0, nz - 1)
    cols = np.clip(np.arange(nxp) - N, 0, nx - 1)
    out = np.zeros((nz, nx))
    np.add.at(out, (rows[:, None].repeat(nxp, 1), cols[None, :].repeat(nzp, 0)), g)
    return out
