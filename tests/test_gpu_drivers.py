"""End-to-end parity of the drop-in drivers on the GPU: the Python facades
(wavefield / migration / fwi, src/Modeling2D.py, src/Migration.py,
src/Inversion.py) and the C++ CLIs (run_modeling / run_migration,
cpp/Modeling.cu, cpp/Migration.cu) read a survey from disk, run through
libpam and write the reference's files; each file is compared with the
oracle run on the same inputs with the same driver semantics."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from oracle.oracle import Grid, Oracle, rel_l2, rtm_onthefly
from pseudoacousticmodeling_amd.survey import parameters
from tests.surveys import write_survey

pytestmark = pytest.mark.gpu

# float32 GPU vs the float32 oracle (relative L2 of whole files)
TOL = 1e-4
BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "pseudoacousticmodeling_amd", "bin")


def _grid(p):
    return Grid(nz=p.nz, nx=p.nx, N=p.N_abc, dz=p.dz, dx=p.dx, dt=p.dt, nt=p.nt, itlag=p.itlag)


def _oracle_setup(orc, p, models, cpp):
    """Grid, expanded models, wavelet and absorbing layer as the chosen
    driver builds them (theta converted from degrees)."""
    g = _grid(p)
    th = None if models.get("theta") is None else np.radians(models["theta"]).astype(np.float32)
    if cpp and th is not None:
        th = (models["theta"] * np.float32(np.pi) / np.float32(180.0)).astype(np.float32)
    m = (models["vp"], models.get("epsilon"), models.get("delta"), th)
    if cpp:
        src = O.ricker_cpp(p.nt, p.dt, p.fcut, np.float32(p.tlag), p.dx, p.dz)
    else:
        src = O.ricker_py(p.nt, p.dt, p.fcut, p.tlag, p.dx, p.dz)
    A = orc.cerjan(p.N_abc, 0.0) if p.ABC == "cerjan" else None
    cpml = O.dampening_const(p.N_abc, p.fcut) if p.ABC == "CPML" else None
    return g, m, src, A, cpml


def _oracle_forward(orc, p, g, m, src, A, cpml, shot, save=False):
    ex = tuple(orc.expand(x, g.N) if x is not None else None for x in m)
    return orc.forward_shot({"acoustic": 0, "VTI": 1, "TTI": 2}[p.approximation], {"cerjan": 0, "CPML": 1}[p.ABC],
                            g, ex, src, int(p.sz[shot]), int(p.sx[shot]), p.rz, p.rx, A=A, cpml=cpml, save=save)


def _facade(path):
    from pseudoacousticmodeling_amd.modeling import wavefield

    p = parameters(path)
    wf = wavefield(p)
    wf.createSourceWavelet()
    wf.initializeWavefields()
    wf.loadModels()
    wf.checkDispersionAndStability()
    return p, wf


@pytest.mark.parametrize("phys,abc", [(0, 0), (1, 1), (2, 0)])
def test_wavefield_files_match_oracle(tmp_path, phys, abc):
    print(f"case wavefield phys={phys} abc={abc}", flush=True)
    path, models = write_survey(str(tmp_path), phys, abc, nshots=2, seed=20 + phys,
                                extra={"snap": True, "step": 60, "last_save": 240})
    p, wf = _facade(path)
    wf.SolveWaveEquation()
    orc = Oracle("f32")
    g, m, src, A, cpml = _oracle_setup(orc, p, models, cpp=False)
    for s in range(p.Nshot):
        f = f"{p.seismogramFolder}seismogram_shot_{s + 1}_Nt{p.nt_data}_Nrec{p.Nrec}_fcut{p.fcut}.bin"
        assert f.endswith(f"seismogram_shot_{s + 1}_Nt301_Nrec19_fcut30.0.bin")
        got = np.fromfile(f, dtype=np.float32).reshape(p.nt_data, p.Nrec)
        ref = _oracle_forward(orc, p, g, m, src, A, cpml, s, save=True)
        e = rel_l2(got, ref["seis"])
        print(f"  shot {s} seismogram rel_l2 {e:.2e}", flush=True)
        assert e <= TOL
        N = p.N_abc
        for k in range(0, 241, 60):
            fs = f"{p.snapshotFolder}{p.approximation}forward_shot_{s + 1}_Nx{p.nx}_Nz{p.nz}_Nt{p.nt}_frame_{k}.bin"
            snap = np.fromfile(fs, dtype=np.float32).reshape(p.nz, p.nx)
            want = ref["save"][k][N:N + p.nz, N:N + p.nx]
            assert rel_l2(snap, want) <= TOL if np.abs(want).max() > 0 else np.abs(snap).max() == 0
        assert not os.path.exists(fs.replace("_frame_240", "_frame_300"))


@pytest.mark.parametrize("phys,abc", [(0, 1), (1, 0)])
def test_migration_image_matches_oracle(tmp_path, phys, abc):
    """RTM through the facade: smoothed migration model (src/utils.py:76-129),
    muted data (src/utils.py:30-49), image / (ilum + 1e-5 max ilum)."""
    print(f"case migration phys={phys} abc={abc}", flush=True)
    from pseudoacousticmodeling_amd.migration import migration

    path, models = write_survey(str(tmp_path), phys, abc, nshots=2, seed=30 + phys, water_rows=4)
    p = parameters(path)
    orc = Oracle("f32")
    g, m, src, A, cpml = _oracle_setup(orc, p, models, cpp=False)
    # observed data: the oracle's gathers of the true model, in the reference's files
    data = []
    for s in range(p.Nshot):
        d = _oracle_forward(orc, p, g, m, src, A, cpml, s)["seis"].astype(np.float32)
        d.tofile(f"{p.seismogramFolder}seismogram_shot_{s + 1}_Nt{p.nt_data}_Nrec{p.Nrec}_fcut{p.fcut}.bin")
        data.append(O.mute(d, p.rec_x, p.rec_z, p.shot_x[s], p.shot_z[s], p.dt, p.shift, p.window, p.v0))
    p, wf = _facade(path)
    mig = migration(wf, p)
    mig.initializeMigrationfields()
    mig.SolveBackwardWaveEquation()
    f = f"{p.imageFolder}migrated_image_{p.approximation}_Nx{p.nx}_Nz{p.nz}.bin"
    assert mig.outputFile == f and os.path.exists(f)
    img = np.fromfile(f, dtype=np.float32).reshape(p.nz, p.nx)
    # the oracle's migration model
    sm = lambda a, par=False: O.smooth_model_py(a, p.sigma, np.abs(a - np.min(a)) < 1e-3, par)  # noqa: E731
    ms = (sm(models["vp"]),) + tuple(sm(models[k], True) if k in models else None for k in ("epsilon", "delta")) \
        + (None,)
    ex = tuple(orc.expand(x, g.N) if x is not None else None for x in ms)
    shots = list(zip(p.sz.tolist(), p.sx.tolist()))
    ref = rtm_onthefly(orc, phys, abc, g, ex, src, shots, p.rz, p.rx, data, A=A, cpml=cpml)
    e = rel_l2(img, ref["image"])
    print(f"  image rel_l2 {e:.2e}", flush=True)
    assert e <= TOL


def test_fwi_iteration_lowers_the_misfit(tmp_path):
    """One L-BFGS iteration of the monoparameter FWI loop from the smoothed
    true model (src/Inversion.py:290-411) lowers the objective and writes the
    reference's gradient, model and history files."""
    print("case fwi", flush=True)
    from pseudoacousticmodeling_amd.inversion import fwi
    from pseudoacousticmodeling_amd.migration import migration

    path, models = write_survey(str(tmp_path), 0, 1, nshots=3, seed=40, water_rows=4,
                                extra={"niter": 2, "sigma": 3.0})
    p, wf = _facade(path)
    wf.SolveWaveEquation()  # observed data
    mig = migration(wf, p)
    mig.initializeMigrationfields()
    hist = str(tmp_path / "outputs" / "history.txt")
    inv = fwi(p, wf, mig, history_file=hist)
    m = inv.solveFullWaveformInversion()
    h = np.loadtxt(hist)
    print(f"  history {h[:, 0].tolist()}", flush=True)
    assert h.shape == (3, 2) and h[0, 0] == 1.0
    assert h[-1, 0] < h[0, 0] and h[-1, 0] < 0.999
    for i in (1, 2):
        assert os.path.exists(f"{p.gradientsFolder}vp_gradient_fwi_iter_{i}_acoustic_Nx{p.nx}_Nz{p.nz}_freq30.0.bin")
        assert os.path.exists(f"{p.estimatedmodelsFolder}fwi_vp_acoustic_Nx{p.nx}_Nz{p.nz}_itr{i}_freq30.0.bin")
    assert os.path.exists(f"{p.modelFolder}fwi_vp_smooth_acoustic_Nx{p.nx}_Nz{p.nz}.bin")
    v = 1.0 / np.sqrt(m)
    assert np.all(v >= p.vmin - 1e-2) and np.all(v <= p.vmax + 1e-2)
    # the water layer is never updated
    np.testing.assert_allclose(v[:4], inv.vp0[:4], rtol=1e-5)


@pytest.mark.parametrize("schedule", ["simultaneous", "hierarchical"])
def test_multiparameter_fwi_lowers_the_misfit(tmp_path, schedule):
    """Two iterations of a multiparameter schedule on a VTI + CPML survey
    (src/Inversion.py:413-940): the misfit falls, the epsilon / delta
    gradients and models of the released parameters are written, and every
    model stays in its range."""
    print(f"case multiparameter fwi {schedule}", flush=True)
    from pseudoacousticmodeling_amd.inversion import fwi
    from pseudoacousticmodeling_amd.migration import migration

    path, models = write_survey(str(tmp_path), 1, 1, nshots=3, seed=41, water_rows=4,
                                extra={"niter": 2, "sigma": 3.0, "multiparameter": True})
    p, wf = _facade(path)
    wf.SolveWaveEquation()  # observed data of the true models
    mig = migration(wf, p)
    mig.initializeMigrationfields()
    hist = str(tmp_path / "outputs" / "history.txt")
    inv = fwi(p, wf, mig, history_file=hist)
    if schedule == "simultaneous":
        out = inv.solveFullWaveformInversion()
    else:
        out = inv.solveFullWaveformInversionMultiparameterHierarchical()
    h = np.loadtxt(hist)
    print(f"  history {h[:, 0].tolist()}", flush=True)
    assert h.shape == (3, 2) and h[-1, 0] < h[0, 0]
    for n in ("vp", "epsilon", "delta"):
        assert os.path.exists(f"{p.estimatedmodelsFolder}fwi_{n}_VTI_Nx{p.nx}_Nz{p.nz}_itr2_freq30.0.bin")
    # epsilon is active from the first of two iterations in both schedules (hierarchical: progress 0.5)
    assert os.path.exists(f"{p.gradientsFolder}epsilon_gradient_fwi_iter_1_VTI_Nx{p.nx}_Nz{p.nz}_freq30.0.bin")
    delta_first = os.path.exists(f"{p.gradientsFolder}delta_gradient_fwi_iter_1_VTI_Nx{p.nx}_Nz{p.nz}_freq30.0.bin")
    assert delta_first == (schedule == "simultaneous")  # hierarchical: delta from progress 0.7
    b = inv._bounds()
    for n, m in out.items():
        assert np.isfinite(m).all()
        assert m.min() >= np.float32(b[n][0]) and m.max() <= np.float32(b[n][1])


def _run(exe, path):
    binary = os.path.join(BIN, exe)
    assert os.path.exists(binary), "CLI not built (python -m pseudoacousticmodeling_amd.build)"
    out = subprocess.run([binary, path], capture_output=True, text=True, timeout=300)
    print(out.stdout[-400:], out.stderr[-400:], flush=True)
    assert out.returncode == 0, out.stderr
    return out


@pytest.mark.parametrize("phys", [0, 1])
def test_cli_modeling_and_migration_match_oracle(tmp_path, phys):
    """run_modeling: seismograms and snapshots with the C++ driver's rules
    (std::round indices, plain Ricker, fcut{:.1f} names, the last snapshot
    written as frame{last_save}, cpp/Modeling.cu:433-470); run_migration:
    separable smoothing with the vp water mask, float32 mute, t = nt-1..0 and
    image / ilum (cpp/Migration.cu:86-236, 359-484), for on-the-fly and
    checkpointed source fields."""
    print(f"case cli phys={phys}", flush=True)
    path, models = write_survey(str(tmp_path), phys, 0, nshots=2, seed=50 + phys, water_rows=4,
                                extra={"snap": True, "step": 50, "last_save": 120})
    p = parameters(path, rounding="round")
    _run("run_modeling", path)
    orc = Oracle("f32")
    g, m, src, A, cpml = _oracle_setup(orc, p, models, cpp=True)
    N = p.N_abc
    data = []
    for s in range(p.Nshot):
        f = f"{p.seismogramFolder}seismogram_shot_{s + 1}_Nt{p.nt_data}_Nrec{p.Nrec}_fcut30.0.bin"
        got = np.fromfile(f, dtype=np.float32).reshape(p.nt_data, p.Nrec)
        ref = _oracle_forward(orc, p, g, m, src, A, cpml, s, save=True)
        e = rel_l2(got, ref["seis"])
        print(f"  shot {s} seismogram rel_l2 {e:.2e}", flush=True)
        assert e <= TOL
        stem = f"{p.snapshotFolder}{p.approximation}forward_shot_{s + 1}Nx{p.nx}_Nz{p.nz}_Nt{p.nt}_frame"
        for k, name in ((50, 50), (100, 120)):
            snap = np.fromfile(f"{stem}{name}.bin", dtype=np.float32).reshape(p.nz, p.nx)
            assert rel_l2(snap, ref["save"][k][N:N + p.nz, N:N + p.nx]) <= TOL
        assert not os.path.exists(f"{stem}100.bin")
        data.append(O.mute_cpp(got, p.rec_x, p.rec_z, p.shot_x[s], p.shot_z[s], p.dt, p.shift, p.window, p.v0))
    _run("run_migration", path)
    f = f"{p.imageFolder}image_{p.approximation}_Nx{p.nx}_Nz{p.nz}.bin"
    img = np.fromfile(f, dtype=np.float32).reshape(p.nz, p.nx)
    mask = np.abs(models["vp"] - models["vp"].min()) < 1e-3
    ms = (O.smooth_model_cpp(models["vp"], p.sigma, mask),) + tuple(
        O.smooth_model_cpp(models[k], p.sigma, mask, True) if k in models else None for k in ("epsilon", "delta")) \
        + (None,)
    ex = tuple(orc.expand(x, g.N) if x is not None else None for x in ms)
    shots = list(zip(p.sz.tolist(), p.sx.tolist()))
    ref = rtm_onthefly(orc, phys, 0, g, ex, src, shots, p.rz, p.rx, data, A=A, t_range="cpp", normalize=2)
    e = rel_l2(img, ref["image"])
    print(f"  image rel_l2 {e:.2e}", flush=True)
    assert e <= TOL
    # checkpointed source field: the same image
    import json

    with open(path) as h:
        cfg = json.load(h)
    cfg.update(migration="checkpoint", step=25)
    with open(path, "w") as h:
        json.dump(cfg, h)
    os.remove(f)
    _run("run_migration", path)
    img2 = np.fromfile(f, dtype=np.float32).reshape(p.nz, p.nx)
    assert rel_l2(img2, img) <= 1e-6
