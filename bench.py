#!/usr/bin/env python3
"""Throughput of the 2D finite-difference engine on one MI355X.

Workload (one "step" = one pass of the hot path over one batch; --workload
picks another single-grid workload for the main leg): forward
modelling of the 64-shot pseudo-acoustic VTI + 50-cell CPML survey
(nz=351, nx=1701, dx=dz=12.5 m -> 451 x 1801 padded cells, nt = 98 + 5001
time steps, 1701 receivers per shot), the shots advanced in batches (the
autotuner's batch, halved per lane, two batches at once on two streams),
seismograms written to HBM.  Inputs (models, wavelet, geometry) are resident
on the device before the timed region.

Prints one JSON line:
  value      Gpt-updates/s = padded cells x time steps x shots / wall time
  roofline   achieved algorithmic HBM GB/s of the dominant kernel (the step
             kernel variant the autotuner picked, at the batch size it picked),
             from HIP events that each launch stamps with the kernel's start
             and end in a profiled run of the same launches, against the
             8 TB/s HBM3E peak; plus the aggregate over the timed runs
  plain_loop the same 64 shots run one shot per launch (the per-shot loop)
  fwi_gradient seconds per shot of the FWI gradient of the same survey, all
             64 shots batched vs the per-shot loop, the source field kept as
             float16 snapshots scaled per step (float32 snapshots and uniform
             checkpoints timed beside it)
  uneven_group sixteen VTI + CPML models of different sizes (201..601 x
             801..2401), 4 shots each, advanced as one group; and their FWI
             gradients, 8 shots each (uneven_group_gradient)
  autograd_gradient the acoustic + CPML workload's FWI gradient through the
             torch.autograd.Function, 50 float16 revolve slots
  linear_adjoint RTM and FWI gradient with the linearised VTI / TTI adjoint
             (Cerjan boundary), 16 shots, 398 steps
  tti_rtm    RTM of the TTI + CPML workload's 32 shots with boundary saving
             (float32 and scaled bfloat16 rims) and with uniform checkpoints,
             illumination-normalised, on a reduced step count
  other_configs acoustic + Cerjan (1 shot, 301 x 501 padded), TTI + CPML (32
             shots, 501 x 2101) and acoustic + CPML (32 shots, 401 x 1301),
             the last two on a reduced step count
  cpu_baseline the CPU restatement of the reference's Numba kernels
             (float32 storage, float64 arithmetic, OpenMP over rows) on a
             bounded sample of the same workload
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PHYS_NAME = {0: "acoustic", 1: "VTI pseudo-acoustic", 2: "TTI pseudo-acoustic"}
METRIC = "Gpt-updates/s & %HBM roofline; FWI gradient s/shot, batched vs plain"


def step_kernel_bytes(w, nshots, fused_psi=True):
    """Algorithmic HBM bytes of one step-kernel launch advancing `nshots` shots
    one time step: per padded cell 12 B per shot (U_k read, U_{k-1} read,
    U_{k+1} write) plus 4 B per model field shared by the shots (acoustic 1,
    VTI 3, TTI 4); per CPML strip cell (counted once per direction, so
    corners twice) 16 B per shot when the kernel updates psi itself (psi read
    and write, zeta read and write; 12 B when a pre-pass wrote psi) plus 8 B
    of shared damping coefficients."""
    nzp, nxp, N = w["nz"] + 2 * w["N"], w["nx"] + 2 * w["N"], w["N"]
    n_m = {0: 1, 1: 3, 2: 4}[w["physics"]]
    b = (12 * nshots + 4 * n_m) * nzp * nxp
    if w["abc"] == 1:
        strip = (nzp - 8) * 2 * (N - 4) + 2 * (N - 4) * (nxp - 8)
        b += ((16 if fused_psi else 12) * nshots + 8) * strip
    return b


def make_engine(w, max_shots):
    from pseudoacousticmodeling_amd.engine import Engine

    eng = Engine(w["nz"], w["nx"], w["N"], w["dz"], w["dx"], w["dt"], w["nt"], w["itlag"],
                 physics=w["physics"], abc=w["abc"], cerjan_factor=w["cerjan_factor"],
                 cpml_d0=w["cpml_d0"], cpml_f_pico=w["cpml_f_pico"], max_shots=max_shots)
    eng.set_model(w["vp"], w["eps"], w["delta"], w["theta"])
    eng.set_wavelet(w["src"])
    eng.set_geometry(w["sz"], w["sx"], w["rz"], w["rx"])
    return eng


def configs_leg(nt_cap):
    """The other single-grid workloads, forward modelling with graph replay
    and the autotuner's choices: acoustic + Cerjan, one shot, every step;
    TTI + CPML, 32 shots, and acoustic + CPML, 32 shots, on the first `nt_cap`
    steps, reported per step as the full runs' rate (labelled reduced)."""
    import torch

    from pseudoacousticmodeling_amd import _lib, workloads

    out = {}
    for build, cap in ((workloads.acoustic_cerjan_1, None), (workloads.tti_cpml_32, nt_cap),
                       (workloads.acoustic_cpml_32, nt_cap)):
        w = build() if cap is None else build(nt_data=cap)
        S = len(w["sx"])
        eng = make_engine(w, S)
        seis = torch.empty((S, w["nt"] - w["itlag"], len(w["rx"])), dtype=torch.float32, device="cuda:0")
        eng.forward(0, S, out=seis)  # tunes, captures
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        eng.forward(0, S, out=seis)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        full = build()
        out[w["name"]] = {"value": workloads.updates(w) / dt / 1e9, "unit": "Gpt-updates/s", "seconds": dt,
                          "steps_run": w["nt"], "steps_full": full["nt"], "reduced": cap is not None,
                          "shots": S, "padded": [w["nz"] + 2 * w["N"], w["nx"] + 2 * w["N"]],
                          "shots_per_launch": eng.get_option(_lib.PAM_OPT_BATCH),
                          "variant": eng.get_option(_lib.PAM_OPT_TILE),
                          "finite": bool(torch.isfinite(seis).all().item())}
        eng.close()
        del seis
        torch.cuda.empty_cache()
    return out


def tti_rtm_leg(nt_data=500, nshots=32):
    """The TTI workload as its configuration states it: RTM of its 32 shots
    (TTI + CPML, 501 x 2101 padded) with the source field rebuilt backwards
    from the rims saved by the forward pass, in float32 and in bfloat16
    scaled per step (boundary saving), and with uniform checkpoints for
    comparison; zero-lag image normalised by the source illumination; the
    first `nt_data` recorded samples (the passes' per-step cost does not
    depend on the step count); the data are the model's own gathers.
    Seconds per shot of the whole RTM (forward pass, rebuild or recompute,
    adjoint pass and imaging)."""
    import torch

    from pseudoacousticmodeling_amd import _lib, workloads

    w = workloads.tti_cpml_32(nshots=nshots, nt_data=nt_data)
    S = len(w["sx"])
    eng = make_engine(w, S)
    data = eng.forward(0, S)
    out = {"unit": "s/shot", "shots": S, "nt": w["nt"], "nt_full": workloads.tti_cpml_32(nshots=1)["nt"],
           "padded": [w["nz"] + 2 * w["N"], w["nx"] + 2 * w["N"]], "normalised": True}
    for name, storage in (("boundary_f32_rims", _lib.PAM_STORE_BOUNDARY_F32),
                          ("boundary_bf16_rims", _lib.PAM_STORE_BOUNDARY_BF16),
                          ("uniform_checkpoints", _lib.PAM_STORE_CHECKPOINT)):
        eng.set_option(_lib.PAM_OPT_STORAGE, storage)
        eng.rtm(data)  # allocates, tunes
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        img, _ = eng.rtm(data)
        torch.cuda.synchronize()
        out[name] = (time.perf_counter() - t0) / S
        out[name + "_finite"] = bool(torch.isfinite(img).all().item())
    eng.close()
    del data
    torch.cuda.empty_cache()
    return out


def autograd_leg(slots=50):
    """The acoustic + CPML workload as an FWI gradient through the
    torch.autograd.Function (pseudoacousticmodeling_amd/autograd.py): 32
    shots, nt = 3001 + itlag, observed gathers of the true model, the misfit
    of the true model Gaussian-smoothed (sigma 250 m) differentiated by
    J.backward() -- the forward modelling, then the exact adjoint with the
    forward state kept in `slots` float16 revolve slots."""
    import numpy as np
    import torch
    from scipy.ndimage import gaussian_filter

    from pseudoacousticmodeling_amd import _lib, workloads
    from pseudoacousticmodeling_amd.autograd import misfit

    w = workloads.acoustic_cpml_32()
    S = len(w["sx"])
    eng = make_engine(w, S)
    dobs = eng.forward(0, S).clone()
    v0 = gaussian_filter(w["vp"].astype(np.float64), 250.0 / w["dx"], mode="nearest").astype(np.float32)
    eng.set_option(_lib.PAM_OPT_STORAGE, _lib.PAM_STORE_REVOLVE_F16)
    eng.set_option(_lib.PAM_OPT_SNAPSHOTS, slots)
    vp = torch.tensor(v0, device="cuda:0", requires_grad=True)
    misfit(vp, eng, dobs).backward()  # tunes, allocates the slots
    torch.cuda.synchronize()
    vp.grad = None
    t0 = time.perf_counter()
    J = misfit(vp, eng, dobs)
    J.backward()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    out = {"unit": "s/shot", "value": dt / S, "seconds": dt, "shots": S, "nt": w["nt"],
           "padded": [w["nz"] + 2 * w["N"], w["nx"] + 2 * w["N"]], "storage": f"revolve, {slots} float16 slots",
           "shots_per_launch": eng.get_option(_lib.PAM_OPT_BATCH) or None, "misfit": float(J.item()),
           "finite": bool(torch.isfinite(vp.grad).all().item())}
    eng.close()
    torch.cuda.empty_cache()
    return out


def linear_adjoint_leg(nt_data=300, nshots=16):
    """RTM and FWI gradient with the linearised VTI / TTI adjoint (the
    reference's adjoint for pseudo-acoustic physics with the Cerjan boundary,
    src/CudaKernels.py:1382-1633): the VTI and TTI workloads' models and
    geometry with a 50-cell Cerjan layer (factor 0.0045) instead of CPML,
    `nshots` shots, the first `nt_data` recorded samples, uniform checkpoints;
    the residual is 1% of the recorded gather.  Seconds per shot."""
    import torch

    from pseudoacousticmodeling_amd import _lib, workloads
    from pseudoacousticmodeling_amd.engine import Engine

    out = {}
    for build in (workloads.vti_cpml_64, workloads.tti_cpml_32):
        w = build(nshots=nshots, nt_data=nt_data)
        S = len(w["sx"])
        eng = Engine(w["nz"], w["nx"], w["N"], w["dz"], w["dx"], w["dt"], w["nt"], w["itlag"], physics=w["physics"],
                     abc=0, cerjan_factor=0.0045, max_shots=S)
        eng.set_model(w["vp"], w["eps"], w["delta"], w["theta"])
        eng.set_wavelet(w["src"])
        eng.set_geometry(w["sz"], w["sx"], w["rz"], w["rx"])
        eng.set_option(_lib.PAM_OPT_STORAGE, _lib.PAM_STORE_CHECKPOINT)
        seis = torch.empty((S, w["nt"] - w["itlag"], len(w["rx"])), dtype=torch.float32, device="cuda:0")
        eng.forward(0, S, out=seis)
        res = seis * 0.01
        rec = {"unit": "s/shot", "shots": S, "nt": w["nt"], "padded": [w["nz"] + 2 * w["N"], w["nx"] + 2 * w["N"]],
               "storage": "uniform checkpoints in HBM"}
        for kind in ("rtm", "gradient"):
            f = eng.rtm if kind == "rtm" else eng.gradient
            f(res)  # allocates, tunes
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            o = f(res)
            torch.cuda.synchronize()
            rec[kind] = (time.perf_counter() - t0) / S
            arr = o["g_m"] if isinstance(o, dict) else o[0]
            rec[kind + "_finite"] = bool(torch.isfinite(arr).all().item())
        out[("vti" if w["physics"] == 1 else "tti") + "_cerjan"] = rec
        eng.close()
        del seis, res
        torch.cuda.empty_cache()
    return out


def group_leg(shots_per_model):
    """Sixteen VTI + CPML models of different sizes (201..601 x 801..2401,
    every pair once), ``shots_per_model`` shots each, nt = 3001 + itlag: all
    shots of all models advanced together (pam_group_forward: cache-sized
    sub-groups, one launch per time step each, on two streams), against the
    same models run one after another (pam_forward each)."""
    import torch

    from pseudoacousticmodeling_amd import workloads
    from pseudoacousticmodeling_amd.engine import Group

    ws = workloads.uneven_vti_group(shots_per_model)
    engs = [make_engine(w, shots_per_model) for w in ws]
    ns = [shots_per_model] * len(engs)
    grp = Group(engs)
    outs = grp.forward(ns)  # tunes, captures
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    grp.forward(ns, outs)
    torch.cuda.synchronize()
    tg = time.perf_counter() - t0
    finite = all(bool(torch.isfinite(o).all().item()) for o in outs)
    for e, o in zip(engs, outs):
        e.forward(0, shots_per_model, out=o)  # tunes, captures
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for e, o in zip(engs, outs):
        e.forward(0, shots_per_model, out=o)
    torch.cuda.synchronize()
    tl = time.perf_counter() - t0
    upd = sum(workloads.updates(w, shots_per_model) for w in ws)
    tile = grp.tile()
    grp.close()
    for e in engs:
        e.close()
    return {"unit": "Gpt-updates/s", "value": upd / tg / 1e9, "per_model_loop": upd / tl / 1e9,
            "speedup": tl / tg, "seconds": tg, "models": len(ws), "shots_per_model": shots_per_model,
            "nt": ws[0]["nt"], "variant": tile, "finite": finite}


def group_gradient_leg(shots_per_model=8):
    """The uneven-grid workload's FWI gradients: sixteen VTI + CPML models of
    different sizes, ``shots_per_model`` shots each, nt = 3001 + itlag,
    residuals = gathers of the model scaled by 1.02 minus the model's own,
    float16-scaled snapshots of every step, float64 gradient norms, one
    pam_group_gradient call (the members in turn, each batched)."""
    import numpy as np
    import torch

    from pseudoacousticmodeling_amd import _lib, workloads
    from pseudoacousticmodeling_amd.engine import Group

    ws = workloads.uneven_vti_group(shots_per_model)
    engs = [make_engine(w, shots_per_model) for w in ws]
    res = []
    for w, e in zip(ws, engs):
        dcal = e.forward(0, shots_per_model)
        e.set_model((w["vp"] * np.float32(1.02)).astype(np.float32), w["eps"], w["delta"], w["theta"])
        res.append(e.forward(0, shots_per_model) - dcal)
        e.set_model(w["vp"], w["eps"], w["delta"], w["theta"])
        e.set_option(_lib.PAM_OPT_STORAGE, _lib.PAM_STORE_ALL_F16)
        del dcal
    grp = Group(engs)
    grp.gradient(res)  # tunes
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    outs, norms = grp.gradient(res)
    torch.cuda.synchronize()
    tg = time.perf_counter() - t0
    n = shots_per_model * len(ws)
    out = {"unit": "s/shot-gradient", "value": tg / n, "seconds": tg, "models": len(ws),
           "shots_per_model": shots_per_model, "shot_gradients": n, "nt": ws[0]["nt"],
           "storage": "float16-scaled snapshots of every step", "gradient_norm2_fp64": norms,
           "finite": all(bool(torch.isfinite(o).all().item()) for o in outs) and all(np.isfinite(norms))}
    grp.close()
    for e in engs:
        e.close()
    del res, outs
    torch.cuda.empty_cache()
    return out


def cpu_baseline(w, seconds):
    """Time the oracle's OpenMP build (float32 storage, float64 arithmetic,
    separate psi / zeta / update passes as the reference's Numba kernels run
    them) on one shot of the same workload for a bounded number of steps."""
    import numpy as np

    from oracle.oracle import Grid, Oracle

    orc = Oracle("mixed")
    g = Grid(nz=w["nz"], nx=w["nx"], N=w["N"], dz=w["dz"], dx=w["dx"], dt=w["dt"], nt=w["nt"], itlag=w["itlag"])
    model = tuple(orc.expand(m, g.N) if m is not None else None for m in (w["vp"], w["eps"], w["delta"], w["theta"]))
    cpml = (w["cpml_d0"], w["cpml_f_pico"]) if w["abc"] == 1 else None
    A = orc.cerjan(g.N, w["cerjan_factor"]) if w["abc"] == 0 else None
    kw = dict(A=A, cpml=cpml)
    sz, sx = int(w["sz"][0]), int(w["sx"][0])
    orc.forward_shot(w["physics"], w["abc"], g, model, w["src"], sz, sx, w["rz"], w["rx"], nsteps=5, **kw)
    t0 = time.perf_counter()
    orc.forward_shot(w["physics"], w["abc"], g, model, w["src"], sz, sx, w["rz"], w["rx"], nsteps=20, **kw)
    per = (time.perf_counter() - t0) / 20
    nsteps = int(max(20, min(g.nt, seconds / max(per, 1e-6))))
    t0 = time.perf_counter()
    orc.forward_shot(w["physics"], w["abc"], g, model, w["src"], sz, sx, w["rz"], w["rx"], nsteps=nsteps, **kw)
    dt = time.perf_counter() - t0
    cells = g.nzp * g.nxp
    threads = int(os.environ.get("OMP_NUM_THREADS", len(os.sched_getaffinity(0))))
    cpu = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": cells * nsteps / dt / 1e9, "unit": "Gpt-updates/s", "cores": threads, "kind": "port",
            "sample": f"1 shot x {nsteps} of {g.nt} time steps of the same workload ({g.nzp}x{g.nxp} padded), "
                      f"{dt:.1f} s on {threads} OpenMP threads of {cpu}; float32 storage, float64 arithmetic",
            "seconds": dt, "nsteps": nsteps, "finite": bool(np.isfinite(model[0]).all())}


def bandwidth_probe():
    """Plain device-memory kernels of known size on 256 MiB float arrays
    (torch's own elementwise kernels): copy (1 read, 1 write), in-place add
    (2 reads, 1 write -- the access pattern of the leapfrog update, which
    writes U_{k+1} over U_{k-1}) and a sum (1 read).  Their FETCH_SIZE /
    WRITE_SIZE counters calibrate the HBM-traffic counters of the step kernel,
    and their rates are the practical bandwidth of this device."""
    import torch

    n = 64 << 20
    a = torch.ones(n, device="cuda:0")
    b = torch.ones(n, device="cuda:0")
    res = {}
    for name, fn, nbytes in (("copy 1R1W", lambda: a.copy_(b), 2 * 4 * n),
                             ("add_ 2R1W", lambda: a.add_(b), 3 * 4 * n),
                             ("sum 1R", lambda: b.sum(), 4 * n)):
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(10):
            fn()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / 10
        res[name] = nbytes / (ms * 1e-3) / 1e9
        print(f"probe {name}: {ms * 1e3:.1f} us per call, {res[name]:.0f} GB/s", flush=True)
    return res


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nt-data", type=int, default=0, help="recorded samples (default: the workload's full record)")
    ap.add_argument("--workload", default="vti_cpml_64", choices=["vti_cpml_64", "tti_cpml_32", "acoustic_cpml_32"],
                    help="the main leg's workload (default: the headline one)")
    ap.add_argument("--tile", type=int, default=-1, help="step-kernel variant (-1: the autotuner's choice)")
    ap.add_argument("--no-plain", action="store_true", help="skip the per-shot loop measurement")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="CPU baseline sample length (0: skip)")
    ap.add_argument("--probe", action="store_true", help="run only the device bandwidth probe")
    ap.add_argument("--no-fwi", action="store_true", help="skip the FWI-gradient legs")
    ap.add_argument("--batch", type=int, default=0,
                    help="shots advanced per launch (0: the autotuner's choice, halved per lane)")
    ap.add_argument("--lanes", type=int, default=2, help="batches advanced at once on their own streams (1 or 2)")
    ap.add_argument("--no-group", action="store_true", help="skip the uneven-grid group leg")
    ap.add_argument("--no-autograd", action="store_true", help="skip the autograd FWI gradient leg")
    ap.add_argument("--no-linear", action="store_true", help="skip the linearised VTI / TTI adjoint leg")
    ap.add_argument("--no-configs", action="store_true", help="skip the other single-grid workloads")
    ap.add_argument("--no-tti-rtm", action="store_true", help="skip the TTI RTM (boundary saving) leg")
    ap.add_argument("--configs-steps", type=int, default=500,
                    help="recorded samples of the reduced TTI + CPML and acoustic + CPML runs")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch the steps one by one instead of replaying captured graphs")
    ap.add_argument("--group-shots", type=int, default=4, help="shots per model in the uneven-grid group leg")
    ap.add_argument("--fwi-plain-shots", type=int, default=8,
                    help="shots timed in the per-shot gradient loop (seconds per shot are reported; 0: skip "
                         "the loop, e.g. to profile the batched gradient alone)")
    args = ap.parse_args(argv)

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(shots shard across GPUs by design, but only the one-GPU path is built and measured)",
              file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from pseudoacousticmodeling_amd import _lib, workloads

    if not torch.cuda.is_available():
        print("bench.py: no GPU visible (torch.cuda.is_available() is False)", file=sys.stderr)
        return 1
    if args.probe:
        bandwidth_probe()
        return 0
    _lib.load()
    build = {"vti_cpml_64": workloads.vti_cpml_64, "tti_cpml_32": workloads.tti_cpml_32,
             "acoustic_cpml_32": workloads.acoustic_cpml_32}[args.workload]
    w = build(nt_data=args.nt_data) if args.nt_data > 0 else build()
    S = len(w["sx"])
    print(f"bench: workload {w['name']} {w['nz'] + 2 * w['N']}x{w['nx'] + 2 * w['N']} padded, nt={w['nt']}, "
          f"{S} shots", file=sys.stderr, flush=True)
    eng = make_engine(w, S)
    if args.tile >= 0:
        eng.set_option(_lib.PAM_OPT_TILE, args.tile)
    if args.batch > 0:
        eng.set_option(_lib.PAM_OPT_BATCH, args.batch)
    if args.no_graph:
        eng.set_option(_lib.PAM_OPT_GRAPH, 0)
    eng.set_option(_lib.PAM_OPT_LANES, args.lanes)
    seis = torch.empty((S, w["nt"] - w["itlag"], len(w["rx"])), dtype=torch.float32, device="cuda:0")
    for _ in range(args.warmup):
        eng.forward(0, S, out=seis)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        eng.forward(0, S, out=seis)
    torch.cuda.synchronize()
    secs = time.perf_counter() - t0
    finite = bool(torch.isfinite(seis).all().item())
    variant = eng.get_option(_lib.PAM_OPT_TILE)
    per_launch = eng.get_option(_lib.PAM_OPT_BATCH)
    upd = workloads.updates(w) * args.steps
    value = upd / secs / 1e9
    # the dominant kernel's average duration: after one more run of the workload
    # (one lane, the same shots per launch), the step kernel launched 200 times
    # back to back on the last batch's fields between two HIP events on its
    # stream (pam_forward with PAM_OPT_PROFILE 2) ...
    eng.set_option(_lib.PAM_OPT_PROFILE, 2)
    eng.forward(0, S, out=seis)
    assert eng.get_option(_lib.PAM_OPT_BATCH) == per_launch
    t = eng.timing()
    kern_ms, launches = t["kernel_ms"], t["kernel_launches"]
    avg_ms = kern_ms / max(launches, 1)
    # ... and one more run of the whole workload with HIP events around every
    # step-kernel launch (launched one by one; each event pair adds the launch's
    # completion signalling to its duration)
    eng.set_option(_lib.PAM_OPT_PROFILE, 1)
    eng.forward(0, S, out=seis)
    t = eng.timing()
    eng.set_option(_lib.PAM_OPT_PROFILE, 0)
    ev_avg_ms = t["kernel_ms"] / max(t["kernel_launches"], 1)
    ev_launches = t["kernel_launches"]
    fused = variant in (4, 5, 6, 10, 11)  # variants that update psi in the step kernel
    bytes_launch = step_kernel_bytes(w, per_launch, fused)
    achieved = bytes_launch / (avg_ms * 1e-3) / 1e9
    print(f"bench: {secs:.3f} s for {args.steps} runs, {per_launch} shots per launch, step kernel variant {variant} "
          f"avg {avg_ms * 1e3:.1f} us over {launches} back-to-back launches ({ev_avg_ms * 1e3:.1f} us with events "
          f"around each of {ev_launches}), seis finite={finite}", file=sys.stderr, flush=True)

    lanes = eng.get_option(_lib.PAM_OPT_LANES)
    n_launch = -(-S // per_launch) * w["nt"] * args.steps
    aggregate = bytes_launch * n_launch / secs / 1e9  # every step launch of the timed runs, wall clock
    plain = None
    if not args.no_plain:
        eng.set_option(_lib.PAM_OPT_LANES, 1)  # one shot at a time, one stream
        eng.set_option(_lib.PAM_OPT_BATCH, 1)
        eng.forward(0, S, out=seis)  # captures the per-shot graphs
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        eng.forward(0, S, out=seis)
        torch.cuda.synchronize()
        ps = time.perf_counter() - t0
        eng.set_option(_lib.PAM_OPT_PROFILE, 1)
        eng.forward(0, 1, out=seis[:1])
        t = eng.timing()
        eng.set_option(_lib.PAM_OPT_PROFILE, 0)
        eng.set_option(_lib.PAM_OPT_BATCH, 0)
        eng.set_option(_lib.PAM_OPT_LANES, lanes)
        pavg = t["kernel_ms"] / max(t["kernel_launches"], 1)
        plain = {"value": workloads.updates(w) / ps / 1e9, "unit": "Gpt-updates/s", "seconds": ps,
                 "k_step_avg_us": pavg * 1e3,
                 "k_step_gbs": step_kernel_bytes(w, 1, eng.get_option(_lib.PAM_OPT_TILE) in (4, 5, 6, 10, 11)) / (pavg * 1e-3) / 1e9,
                 "batched_speedup": ps / (secs / args.steps)}
        print(f"bench: plain per-shot loop {ps:.3f} s", file=sys.stderr, flush=True)
    fwi = None
    if not args.no_fwi:
        # FWI gradient of the same survey: adjoint-state gradient w.r.t. m = 1/vp^2 of all
        # 64 shots; the residual is a recorded gather of the same shape.  The source field
        # of every step is kept as float16 scaled per step and shot (half of float32
        # save-all; the whole survey at once would need 390 GB, so the batches keep it in
        # turn), the gradient within 1e-4 of the float64 oracle; float32 snapshots and
        # uniform checkpoints are timed beside it
        storages = (("float16-scaled snapshots of every step", _lib.PAM_STORE_ALL_F16),
                    ("float32 snapshots of every step", _lib.PAM_STORE_ALL_F32),
                    ("uniform checkpoints in HBM", _lib.PAM_STORE_CHECKPOINT))
        res = seis
        np_ = min(args.fwi_plain_shots, S)
        recs = []
        for sname, sopt in storages:
            eng.set_option(_lib.PAM_OPT_STORAGE, sopt)
            eng.set_option(_lib.PAM_OPT_BATCH, args.batch)
            eng.set_option(_lib.PAM_OPT_LANES, lanes)
            eng.gradient(res)  # allocates snapshots or checkpoints, tunes
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            out_s = eng.gradient(res)
            torch.cuda.synchronize()
            tb = time.perf_counter() - t0
            sfinite = bool(torch.isfinite(out_s["g_m"]).all().item())
            del out_s
            tp = None
            if np_ > 0:
                # the per-shot loop: one shot at a time on one stream, the same storage
                eng.set_option(_lib.PAM_OPT_BATCH, 1)
                eng.set_option(_lib.PAM_OPT_LANES, 1)
                eng.gradient(res[:1])
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                eng.gradient(res[:np_])
                torch.cuda.synchronize()
                tp = time.perf_counter() - t0
            recs.append({"storage": sname, "batched": tb / S, "plain": tp / np_ if tp else None,
                         "speedup": (tp / np_) / (tb / S) if tp else None, "finite": sfinite})
            torch.cuda.empty_cache()
        eng.set_option(_lib.PAM_OPT_BATCH, args.batch)
        eng.set_option(_lib.PAM_OPT_LANES, lanes)
        head = recs[0]
        gb, gp, gfinite = head["batched"] * S, (head["plain"] * np_ if head["plain"] else None), head["finite"]
        fwi = {"unit": "s/shot", "batched": head["batched"], "plain": head["plain"], "speedup": head["speedup"],
               "shots_batched": S, "shots_plain_loop": np_, "storage": head["storage"], "lanes": lanes,
               "finite": gfinite, "other_storages": {r["storage"]: {k: r[k] for k in ("batched", "plain", "speedup",
                                                                                       "finite")}
                                                     for r in recs[1:]}}
        plain_txt = f", {gp / np_ * 1e3:.1f} ms/shot plain" if gp else ""
        print(f"bench: FWI gradient {gb / S * 1e3:.1f} ms/shot batched{plain_txt}", file=sys.stderr, flush=True)
    eng.close()
    del seis
    torch.cuda.empty_cache()

    group = None
    if not args.no_group:
        group = group_leg(args.group_shots)
        print(f"bench: uneven-grid group {group['value']:.1f} Gpt-updates/s together, "
              f"{group['per_model_loop']:.1f} model by model", file=sys.stderr, flush=True)

    group_grad = None
    if not args.no_group:
        group_grad = group_gradient_leg()
        print(f"bench: uneven-grid group FWI gradients {group_grad['value'] * 1e3:.1f} ms/shot-gradient "
              f"({group_grad['shot_gradients']} shot-gradients, {group_grad['seconds']:.2f} s)",
              file=sys.stderr, flush=True)

    autograd = None
    if not args.no_autograd:
        autograd = autograd_leg()
        print(f"bench: autograd FWI gradient (acoustic CPML, revolve 50 float16 slots) "
              f"{autograd['value'] * 1e3:.1f} ms/shot", file=sys.stderr, flush=True)

    linear = None
    if not args.no_linear:
        linear = linear_adjoint_leg()
        for k, v in linear.items():
            print(f"bench: linearised adjoint {k}: RTM {v['rtm'] * 1e3:.2f} ms/shot, FWI gradient "
                  f"{v['gradient'] * 1e3:.2f} ms/shot ({v['shots']} shots, nt={v['nt']})", file=sys.stderr, flush=True)

    tti_rtm = None
    if not args.no_tti_rtm:
        tti_rtm = tti_rtm_leg()
        print(f"bench: TTI RTM (32 shots, nt={tti_rtm['nt']}): float32 rims {tti_rtm['boundary_f32_rims'] * 1e3:.1f}, "
              f"bfloat16 rims {tti_rtm['boundary_bf16_rims'] * 1e3:.1f}, uniform checkpoints "
              f"{tti_rtm['uniform_checkpoints'] * 1e3:.1f} ms/shot", file=sys.stderr, flush=True)

    others = None
    if not args.no_configs:
        others = configs_leg(args.configs_steps)
        for k, v in others.items():
            print(f"bench: {k} {v['value']:.1f} Gpt-updates/s ({v['steps_run']} of {v['steps_full']} steps, "
                  f"{v['shots']} shots)", file=sys.stderr, flush=True)

    cpu = None
    if args.cpu_seconds > 0:
        print("bench: cpu baseline ...", file=sys.stderr, flush=True)
        cpu = cpu_baseline(w, args.cpu_seconds)
        print(f"bench: cpu {cpu['value']:.3f} Gpt-updates/s ({cpu['sample']})", file=sys.stderr, flush=True)

    # HBM traffic per launch of this variant at this batch, from rocprofv3
    # FETCH_SIZE / WRITE_SIZE runs of the same workload kept in profiles/ (not
    # measured in this run: counters need a profiler run of their own)
    traffic, traffic_note = None, "no committed counter run for this variant and batch"
    tpath = os.path.join(ROOT, "profiles", f"traffic_{w['name']}.json")
    if os.path.exists(tpath) and args.nt_data == 0:
        for rec in json.load(open(tpath)).get("launches", []):
            if rec.get("variant") == variant and rec.get("shots_per_launch") == per_launch:
                traffic = rec["traffic_bytes"]
                traffic_note = ("from the committed rocprofv3 counter runs in %s (FETCH_SIZE x %.1f + WRITE_SIZE x "
                                "%.1f), not measured in this run" % (rec["source"], rec["fetch_scale"],
                                                                     rec["write_scale"]))
    out = {
        "metric": METRIC, "value": value, "unit": "Gpt-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": secs / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{w['name']}: {PHYS_NAME[w['physics']]} forward modelling + 50-cell CPML, "
                               f"{w['nz'] + 2 * w['N']}x{w['nx'] + 2 * w['N']} padded cells, nt={w['nt']}, "
                               f"{S} shots, {len(w['rx'])} receivers per shot",
                   "global_batch": S, "nt": w["nt"], "shots_per_launch": per_launch, "lanes": lanes,
                   "parallelism": f"1 GPU, {per_launch} shots per launch, {lanes} batches at once on {lanes} streams"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_note": traffic_note,
                     "kernel": "step kernel variant %d" % variant,
                     "bytes_per_launch": bytes_launch, "avg_launch_us": avg_ms * 1e3,
                     "avg_launch_note": "the step kernel launched 200 times back to back on one batch's fields "
                                        "after its run, between two HIP events on its stream (no CPML pre-pass "
                                        "between the launches)",
                     "avg_launch_us_events_each": ev_avg_ms * 1e3,
                     "events_each_note": "every step launch of one more run of the workload bracketed by its own "
                                         "HIP events (one lane); the event pair adds the launch's completion "
                                         "signalling to each duration",
                     "aggregate_GBs": aggregate, "aggregate_frac": aggregate / HBM_PEAK_GBS,
                     "aggregate_note": "algorithmic bytes of every step launch of the timed runs / their wall "
                                       "time (lanes overlapping, CPML pre-pass included)"},
        "plain_loop": plain,
        "fwi_gradient": fwi,
        "uneven_group": group,
        "autograd_gradient": autograd,
        "linear_adjoint": linear,
        "uneven_group_gradient": group_grad,
        "tti_rtm": tti_rtm,
        "other_configs": others,
        "cpu_baseline": cpu,
        "finite": finite,
    }
    if cpu is not None:
        out["cpu_baseline"] = {k: cpu[k] for k in ("value", "unit", "cores", "kind", "sample")}
    print(json.dumps(out), flush=True)
    return 0 if finite else 1


if __name__ == "__main__":
    sys.exit(main())
