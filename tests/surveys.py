"""On-disk surveys for the driver tests: a Parameters.json with the
reference's key set (inputs/Parameters.json), the geometry CSVs
(index,coordx,coordz) and raw float32 [nz][nx] model files in a temporary
folder, sized so the CPU oracle runs each shot in about a second."""
import json
import os

import numpy as np

from tests.problems import smooth_field

APPROX = {0: "acoustic", 1: "VTI", 2: "TTI"}
ABC = {0: "cerjan", 1: "CPML"}


def write_csv(path, x, z):
    with open(path, "w") as f:
        f.write("index,coordx,coordz\n")
        for i, (a, b) in enumerate(zip(x, z)):
            f.write(f"{i},{a},{b}\n")


def write_survey(root, phys=0, abc=0, nshots=2, seed=5, L=550.0, D=390.0, T=0.3, dx=10.0, dt=0.001, fcut=30.0,
                 N_abc=12, sigma=2.0, migration="onthefly", extra=None, water_rows=0):
    """Writes models, geometry and Parameters.json under ``root``; returns
    (path of Parameters.json, dict of the interior models as written)."""
    rng = np.random.default_rng(seed)
    nx = int(round(L / dx)) + 1
    nz = int(round(D / dx)) + 1
    for sub in ("models", "seismograms", "images", "snapshots", "gradients", "estimated", "checkpoints", "outputs"):
        os.makedirs(os.path.join(root, sub), exist_ok=True)
    vp = smooth_field(rng, nz, nx, 1800.0, 3000.0)
    if water_rows:
        vp[:water_rows] = 1500.0
    models = {"vp": vp}
    if phys >= 1:
        models["epsilon"] = smooth_field(rng, nz, nx, 0.0, 0.25)
        models["delta"] = np.minimum(smooth_field(rng, nz, nx, 0.0, 0.1), models["epsilon"]).astype(np.float32)
    if phys == 2:
        models["theta"] = smooth_field(rng, nz, nx, -30.0, 30.0)  # degrees on disk
    files = {}
    for k, m in models.items():
        f = os.path.join(root, "models", f"{k}.bin")
        m.astype(np.float32).tofile(f)
        files[k] = f
    sx = 100.0 + 300.0 * np.arange(nshots) / max(nshots - 1, 1)
    sz = np.full(nshots, 20.0)
    rx = np.arange(0.0, L + 1e-9, 30.0)
    rz = np.full(rx.shape, 20.0)
    write_csv(os.path.join(root, "rec.csv"), rx, rz)
    write_csv(os.path.join(root, "src.csv"), sx, sz)
    p = {
        "unit": "GPU", "approximation": APPROX[phys], "migration": migration, "ABC": ABC[abc],
        "dx": dx, "dz": dx, "dt": dt, "L": L, "D": D, "T": T, "N_abc": N_abc, "fcut": fcut,
        "snap": False, "step": 0, "last_save": 0,
        "seismogramFolder": os.path.join(root, "seismograms") + "/",
        "imageFolder": os.path.join(root, "images") + "/",
        "snapshotFolder": os.path.join(root, "snapshots") + "/",
        "modelFolder": os.path.join(root, "models") + "/",
        "checkpointFolder": os.path.join(root, "checkpoints") + "/",
        "estimatedmodelsFolder": os.path.join(root, "estimated") + "/",
        "gradientsFolder": os.path.join(root, "gradients") + "/",
        "rec_file": os.path.join(root, "rec.csv"), "src_file": os.path.join(root, "src.csv"),
        "vpFile": files["vp"], "epsilonFile": files.get("epsilon", ""), "deltaFile": files.get("delta", ""),
        "thetaFile": files.get("theta", ""),
        "layer2": False, "layer3": False, "gradientmodel": False, "diffractor": False, "modelfromvp": False,
        "waterlayer": False, "idx_water": 0, "shift": 0.04, "window": 0.05, "v0": 1500.0, "sigma": sigma,
        "dvel": 0.0, "ratio": 0.0, "mirror": False, "reciprocity": False, "niter": 1, "fwi": False,
        "freqs": [fcut], "vmax": 3200.0, "vmin": 1500.0, "epsmin": 0.0, "epsmax": 0.3, "deltamin": 0.0,
        "deltamax": 0.15, "thetamin": -40.0, "thetamax": 40.0, "multiparameter": False,
    }
    if extra:
        p.update(extra)
    path = os.path.join(root, "Parameters.json")
    with open(path, "w") as f:
        json.dump(p, f, indent=2)
    return path, models
