"""Seeded synthetic problems shared by the tests (small grids the CPU oracle
finishes in seconds)."""
import numpy as np

from oracle.oracle import Grid, ricker_cpp, dampening_const


def smooth_field(rng, nz, nx, lo, hi, passes=6):
    f = rng.standard_normal((nz, nx))
    for _ in range(passes):
        f = 0.25 * (np.roll(f, 1, 0) + np.roll(f, -1, 0) + np.roll(f, 1, 1) + np.roll(f, -1, 1))
    f = (f - f.min()) / max(f.max() - f.min(), 1e-12)
    return (lo + (hi - lo) * f).astype(np.float32)


def make_problem(phys=0, abc=0, nz=40, nx=56, N=12, nt=360, itlag=30, nshots=1, seed=1, amp=1.0,
                 dx=10.0, dt=0.001, theta_deg=30.0):
    rng = np.random.default_rng(seed)
    g = Grid(nz=nz, nx=nx, N=N, dz=dx, dx=dx, dt=dt, nt=nt, itlag=itlag)
    vp = smooth_field(rng, nz, nx, 1500.0, 3000.0)
    eps = dl = th = None
    if phys >= 1:
        eps = smooth_field(rng, nz, nx, 0.0, 0.25)
        dl = np.minimum(smooth_field(rng, nz, nx, 0.0, 0.1), eps).astype(np.float32)
    if phys == 2:
        th = np.deg2rad(smooth_field(rng, nz, nx, -theta_deg, theta_deg)).astype(np.float32)
    fcut = 30.0
    tlag = 2.0 * np.sqrt(np.pi) / fcut
    src = (ricker_cpp(nt, dt, fcut, tlag, dx, dx) * np.float32(amp)).astype(np.float32)
    sz = np.array([N + 4 + (s % 3) for s in range(nshots)], dtype=np.int32)
    sx = np.array([N + 6 + (s * 7) % max(nx - 12, 1) for s in range(nshots)], dtype=np.int32)
    rx = np.arange(N, N + nx, 3, dtype=np.int32)
    rz = np.full_like(rx, N + 2)
    d0, fp = dampening_const(N, fcut)
    return dict(g=g, vp=vp, eps=eps, dl=dl, th=th, src=src, sz=sz, sx=sx, rz=rz, rx=rx, phys=phys, abc=abc,
                d0=d0, fp=fp, fcut=fcut)


def oracle_forward(orc, pb, shot=0, cerjan_factor=0.0):
    g = pb["g"]
    model = tuple(orc.expand(m, g.N) if m is not None else None for m in (pb["vp"], pb["eps"], pb["dl"], pb["th"]))
    A = orc.cerjan(g.N, cerjan_factor) if pb["abc"] == 0 else None
    cpml = (pb["d0"], pb["fp"]) if pb["abc"] == 1 else None
    return orc.forward_shot(pb["phys"], pb["abc"], g, model, pb["src"], int(pb["sz"][shot]), int(pb["sx"][shot]),
                            pb["rz"], pb["rx"], A=A, cpml=cpml)


def make_engine(pb, max_shots=1, cerjan_factor=0.0):
    from pseudoacousticmodeling_amd.engine import Engine

    g = pb["g"]
    eng = Engine(g.nz, g.nx, g.N, g.dz, g.dx, g.dt, g.nt, g.itlag, physics=pb["phys"], abc=pb["abc"],
                 cerjan_factor=cerjan_factor, cpml_d0=pb["d0"], cpml_f_pico=pb["fp"], max_shots=max_shots)
    eng.set_model(pb["vp"], pb["eps"], pb["dl"], pb["th"])
    eng.set_wavelet(pb["src"])
    eng.set_geometry(pb["sz"], pb["sx"], pb["rz"], pb["rx"])
    return eng
