"""Parameters.json and acquisition-geometry reader: the `parameters` object of
src/survey.py:5-141 with the same attribute names, so driver code written
against the reference reads the same fields.

Derived sizes follow the Python driver (int() truncation, src/survey.py:37-47)
by default; ``rounding="round"`` gives the C++ driver's std::round rules
(cpp/Survey.cpp:101-108, 150-158).  The two differ, e.g. itlag = 148 vs 147 for
fcut = 24 Hz at dt = 1 ms.
"""
from __future__ import annotations

import csv
import json

import numpy as np

from . import prep


def read_geometry_csv(path):
    """(coordx, coordz) columns of a geometry CSV with header
    ``index,coordx,coordz`` (src/createGeometry.py:27-32); the C++ reader
    takes columns 2-3 (cpp/Survey.cpp:20-38)."""
    xs, zs = [], []
    with open(path, newline="") as f:
        rows = list(csv.reader(f))
    header = [h.strip() for h in rows[0]]
    if "coordx" in header and "coordz" in header:
        ix, iz = header.index("coordx"), header.index("coordz")
        body = rows[1:]
    else:  # no header: index, x, z
        ix, iz = 1, 2
        body = rows
    for r in body:
        if not r or not r[0].strip():
            continue
        xs.append(float(r[ix]))
        zs.append(float(r[iz]))
    return np.asarray(xs, dtype=np.float64), np.asarray(zs, dtype=np.float64)


class parameters:
    """All keys of inputs/Parameters.json plus the derived sizes and the
    source / receiver grid indices (padded grid)."""

    def __init__(self, parameters_path, rounding="trunc"):
        self.parameters_path = parameters_path
        self.rounding = rounding
        self.readParameters()
        self.readAcquisitionGeometry()

    def readParameters(self):
        with open(self.parameters_path) as f:
            p = json.load(f)
        self.parameters = p
        self.unit = p.get("unit", "GPU")
        self.approximation = p["approximation"]
        self.migration = p.get("migration", "onthefly")
        self.ABC = p["ABC"]
        self.dx, self.dz, self.dt = p["dx"], p["dz"], p["dt"]
        self.L, self.D, self.T = p["L"], p["D"], p["T"]
        self.N_abc = p["N_abc"]
        self.fcut = p["fcut"]
        self.update_sizes()
        for key in ("seismogramFolder", "imageFolder", "snapshotFolder", "modelFolder", "checkpointFolder",
                    "estimatedmodelsFolder", "gradientsFolder", "rec_file", "src_file", "vpFile", "thetaFile",
                    "epsilonFile", "deltaFile"):
            setattr(self, key, p.get(key, ""))
        self.snap = p.get("snap", False)
        self.step = p.get("step", 0)
        self.last_save = p.get("last_save", 0)
        for key, default in (("layer2", False), ("layer3", False), ("gradientmodel", False), ("diffractor", False),
                             ("modelfromvp", False), ("waterlayer", False), ("idx_water", 0), ("shift", 0.0),
                             ("window", 0.0), ("v0", 1500.0), ("sigma", 0.0), ("dvel", 0.0), ("ratio", 0.0),
                             ("mirror", False), ("reciprocity", False), ("niter", 0), ("fwi", False),
                             ("freqs", []), ("vmax", 0.0), ("vmin", 0.0), ("epsmin", 0.0), ("epsmax", 0.0),
                             ("deltamin", 0.0), ("deltamax", 0.0), ("thetamin", 0.0), ("thetamax", 0.0),
                             ("multiparameter", False)):
            setattr(self, key, p.get(key, default))
        # optional keys of this build (absent from the reference's file)
        self.cerjan_factor = p.get("cerjan_factor", 0.0)
        self.wavelet = p.get("wavelet", "python")   # "python": Ricker + sqrt(i w); "cpp": plain Ricker
        self.storage = p.get("storage", "float32")  # snapshot precision: float32 | float16 | bfloat16

    def update_sizes(self):
        """tlag, itlag, nt_data, nt, nx, nz and the padded sizes from the
        current dx, dz, dt, L, D, T, fcut, N_abc."""
        s = prep.derived_sizes(self.L, self.D, self.T, self.dx, self.dz, self.dt, self.fcut, self.rounding)
        self.tlag = s["tlag"]
        self.itlag = s["itlag"]
        self.nt_data = s["nt_data"]
        self.nt = s["nt"]
        self.nx = s["nx"]
        self.nz = s["nz"]
        self.nx_abc = self.nx + 2 * self.N_abc
        self.nz_abc = self.nz + 2 * self.N_abc
        self.x = np.arange(self.nx) * self.dx
        self.z = np.arange(self.nz) * self.dz
        self.t = np.arange(self.nt) * self.dt

    def readAcquisitionGeometry(self):
        """Receiver / source coordinates, reciprocity swap and the mirror
        shift of src/survey.py:114-141."""
        self.rec_x, self.rec_z = read_geometry_csv(self.rec_file)
        self.shot_x, self.shot_z = read_geometry_csv(self.src_file)
        if self.reciprocity:
            self.shot_x, self.rec_x = self.rec_x.copy(), self.shot_x.copy()
            self.shot_z, self.rec_z = self.rec_z.copy(), self.shot_z.copy()
        self.rx = prep.grid_index(self.rec_x, self.dx, self.N_abc, self.rounding)
        self.rz = prep.grid_index(self.rec_z, self.dz, self.N_abc, self.rounding)
        self.sx = prep.grid_index(self.shot_x, self.dx, self.N_abc, self.rounding)
        self.sz = prep.grid_index(self.shot_z, self.dz, self.N_abc, self.rounding)
        if self.mirror:
            self.rz = self.rz - np.int32(self.idx_water)
            self.sz = self.sz - np.int32(self.idx_water)
        self.Nrec = len(self.rec_x)
        self.Nshot = len(self.shot_x)
