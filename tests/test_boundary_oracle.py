"""Pins of the oracle's boundary-saving restatement (src/Migration.py:1266-1410
semantics, time alignment re-derived): in float64 the source field rebuilt
backwards in time from the saved rims must give the same RTM image and FWI
gradient as the source field kept at every step, wherever the adjoint reads S
only on the interior; a plausible mistake planted in the rebuild (the source
sample of the wrong step, no source correction, rims of the wrong step) must
break that equality; and the float16 / bfloat16 rim quantisers must agree
with numpy's float16 and torch's bfloat16 conversions."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.problems import make_problem
from tests.test_gpu_adjoint import _data


def _run(pb, data, fwi, storage, rim_kind="f32", kind="f64", multiparameter=False, t_range="py"):
    orc = O.Oracle(kind)
    g = pb["g"]
    model = tuple(orc.expand(m, g.N) if m is not None else None for m in (pb["vp"], pb["eps"], pb["dl"], pb["th"]))
    A = orc.cerjan(g.N) if pb["abc"] == 0 else None
    cpml = (pb["d0"], pb["fp"]) if pb["abc"] == 1 else None
    shots = list(zip(pb["sz"].tolist(), pb["sx"].tolist()))
    interior = (pb["vp"], pb["eps"], pb["dl"], pb["th"])
    return O.rtm_onthefly(orc, pb["phys"], pb["abc"], g, model, pb["src"], shots, pb["rz"], pb["rx"], data, A=A,
                          cpml=cpml, fwi=fwi, multiparameter=multiparameter, model_interior=interior,
                          storage=storage, rim_kind=rim_kind, t_range=t_range)


# cases whose adjoint reads S only on the interior (pressure-form adjoint):
# acoustic with either boundary, VTI / TTI with CPML
@pytest.mark.parametrize("phys,abc", [(0, 0), (0, 1), (1, 1), (2, 1)])
@pytest.mark.parametrize("fwi", [False, True])
def test_rebuild_equals_save_all_in_float64(phys, abc, fwi):
    pb = make_problem(phys, abc, nshots=1, seed=300 + 2 * phys + abc, nt=160)
    data = _data(pb, 1, seed=30)
    multi = fwi and phys >= 1
    ref = _run(pb, data, fwi, "all", multiparameter=multi)
    got = _run(pb, data, fwi, "boundary", multiparameter=multi)
    keys = ["image", "ilum"] + (["g_eps", "g_delta"] if multi else [])
    for k in keys:
        err = O.rel_l2(got[k], ref[k])
        assert err < 1e-9, f"{k}: {err:.3e}"


def test_rebuild_equals_save_all_over_the_cpp_range():
    pb = make_problem(0, 1, nshots=1, seed=311, nt=140)
    data = _data(pb, 1, seed=31)
    ref = _run(pb, data, False, "all", t_range="cpp")
    got = _run(pb, data, False, "boundary", t_range="cpp")
    assert O.rel_l2(got["image"], ref["image"]) < 1e-9


@pytest.mark.parametrize("mistake", ["wavelet_index", "no_source", "rim_step"])
def test_planted_mistakes_in_the_rebuild_are_caught(monkeypatch, mistake):
    """The source (make_problem puts it 4-6 cells inside the interior edge)
    lies in the region the rebuild recomputes, so its correction matters."""
    pb = make_problem(0, 0, nshots=1, seed=320, nt=160)
    data = _data(pb, 1, seed=32)
    ref = _run(pb, data, True, "all")

    def broken_step(self, need_d=False):
        g, sl, lev = self.g, self.sl, self.lev
        vp, eps, dl, th = self.model
        hi_int = self.hi[sl].copy()
        Uf, Uc = self.hi, self.lo
        self.orc.update(self.phys, Uf, Uc, g, vp, eps, dl, th)
        keep = Uf[sl].copy()
        Uf[...] = 0
        Uf[sl] = keep
        if mistake == "wavelet_index":
            Uf[self.sz, self.sx] += self.src[lev]
        elif mistake == "rim_step":
            Uf[self.sz, self.sx] += self.src[lev + 1] if lev + 1 < g.nt else 0.0
        Uf[sl][self.mask] = self.rims[lev] if mistake == "rim_step" else self.rims[lev - 1]
        d = (Uf[sl] - 2.0 * Uc[sl]) + hi_int
        self.hi, self.lo = Uc, Uf
        self.lev = lev - 1
        return d

    monkeypatch.setattr(O._BoundaryRebuild, "_step", broken_step)
    got = _run(pb, data, True, "boundary")
    assert O.rel_l2(got["image"], ref["image"]) > 1e-4


def test_bfloat16_rounding_matches_torch():
    import torch

    rng = np.random.default_rng(5)
    x = np.concatenate([rng.standard_normal(4000).astype(np.float32) * 10.0 ** rng.integers(-8, 8, 4000),
                        np.float32([1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -1.0 - 2.0 ** -8, 0.0, 65504.0])])
    x = x.astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).float().numpy()
    assert np.array_equal(O._bf16_round(x), ref)
    # ties to even, by hand: 1 + 2^-8 -> 1, 1 + 3 2^-8 -> 1 + 2^-6
    assert O._bf16_round(np.float32([1.0 + 2.0 ** -8]))[0] == 1.0
    assert O._bf16_round(np.float32([1.0 + 3 * 2.0 ** -8]))[0] == np.float32(1.0 + 2.0 ** -6)


@pytest.mark.parametrize("kind,bits", [("f16", 11), ("bf16", 8)])
def test_scaled_quantiser_error_bound(kind, bits):
    """v / m rounded to a p-bit significand and scaled back: each value within
    2^-p m of itself (values near m lose at most half an ulp of 1; smaller
    ones less, until float16's subnormal range, which 2^-p m still bounds)."""
    rng = np.random.default_rng(6)
    v = (rng.standard_normal(20000) * 3e-7).astype(np.float32)
    q = O.quantize_scaled(v, kind)
    m = np.max(np.abs(v))
    assert np.max(np.abs(q - v)) <= 2.0 ** -bits * m * (1 + 1e-6)
    assert np.array_equal(O.quantize_scaled(v, "f32"), v)
    f16 = (v * (np.float32(1.0) / np.float32(m))).astype(np.float16).astype(np.float32) * np.float32(m)
    if kind == "f16":
        assert np.array_equal(q, f16)


def test_rim_mask_is_the_four_cell_frame():
    m = O.rim_mask(12, 15)
    assert m.sum() == 12 * 15 - 4 * 7
    assert m[:4].all() and m[-4:].all() and m[:, :4].all() and m[:, -4:].all()
    assert not m[4:-4, 4:-4].any()


# ---------------------------------------------------------------- random boundary
def test_random_rebuild_reverses_the_undamped_forward_in_float64():
    """The whole-grid rebuild of the random-boundary strategy reproduces every
    step of the undamped forward pass (float64, to 1e-9), with the source
    inside the grid's rebuilt region: it pins the time alignment and the
    source correction."""
    from pseudoacousticmodeling_amd import prep

    pb = make_problem(1, 0, nshots=1, seed=340, nt=150)
    g = pb["g"]
    orc = O.Oracle("f64")
    vp = prep.random_boundary(orc.expand(pb["vp"], g.N), g.N, g.dx, g.dz, 300.0, 40.0, np.random.default_rng(3))
    model = (orc.arr(vp), orc.expand(pb["eps"], g.N), orc.expand(pb["dl"], g.N), None)
    sz, sx = int(pb["sz"][0]), int(pb["sx"][0])
    fw = orc.forward_shot(1, 0, g, model, pb["src"], sz, sx, pb["rz"], pb["rx"], A=np.ones(g.N), save=True)
    S = fw["save"]
    rb = O._RandomRebuild(orc, 1, g, model, S, fw["current"], orc.arr(pb["src"]), sz, sx)
    scale = np.abs(S).max()
    for t in range(g.nt - 2, 0, -1):
        St, Dt = rb.get(t, True)
        assert np.abs(St - S[t]).max() <= 1e-9 * scale, f"S_{t}"
        sl = rb.sl
        D_ref = S[t + 1][sl] - 2.0 * S[t][sl] + S[t - 1][sl]
        assert np.abs(Dt - D_ref).max() <= 1e-9 * scale, f"D_{t}"


def test_random_boundary_model_statistics():
    """Properties of prep.random_boundary (src/Migration.py:382-458): the
    layer's cells differ from the edge velocity, more so towards the outer
    edge (the spread ramps up linearly); every value lies in
    [vmin - dvel, vmax + dvel]; the interior moves only by the blobs' tails;
    two seeds give different layers and one seed the same."""
    from pseudoacousticmodeling_amd import prep

    N, nz, nx, dvel = 20, 60, 90, 300.0
    vp = np.pad(np.full((nz, nx), 2000.0, np.float32), N, mode="edge")
    a = prep.random_boundary(vp, N, 10.0, 10.0, dvel, 60.0, np.random.default_rng(7))
    b = prep.random_boundary(vp, N, 10.0, 10.0, dvel, 60.0, np.random.default_rng(7))
    c = prep.random_boundary(vp, N, 10.0, 10.0, dvel, 60.0, np.random.default_rng(8))
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert a.min() >= 2000.0 - dvel - 1e-3 and a.max() <= 2000.0 + dvel + 1e-3
    inner = a[N:N + nz, N:N + nx]
    assert np.abs(inner - 2000.0).max() < dvel
    # deviation grows towards the outer edge of the left layer (mean over rows)
    dev = np.abs(a[N:N + nz, :N] - 2000.0).mean(axis=0)
    assert dev[:5].mean() > 2.0 * dev[-5:].mean()


@pytest.mark.parametrize("mistake", ["no_source", "wavelet_index", "source_sign"])
def test_planted_mistakes_in_the_random_rebuild_are_caught(monkeypatch, mistake):
    """A source correction dropped, taken one sample early or with the wrong
    sign breaks the float64 reversal of the undamped forward pass."""
    from pseudoacousticmodeling_amd import prep

    pb = make_problem(0, 0, nshots=1, seed=341, nt=150)
    g = pb["g"]
    orc = O.Oracle("f64")
    vp = prep.random_boundary(orc.expand(pb["vp"], g.N), g.N, g.dx, g.dz, 300.0, 40.0, np.random.default_rng(4))
    model = (orc.arr(vp), None, None, None)
    sz, sx = int(pb["sz"][0]), int(pb["sx"][0])
    fw = orc.forward_shot(0, 0, g, model, pb["src"], sz, sx, pb["rz"], pb["rx"], A=np.ones(g.N), save=True)
    S = fw["save"]

    def broken_step(self, need_d=False):
        lev = self.lev
        vp_, eps, dl, th = self.model
        Uf, Uc = self.hi, self.lo
        self.orc.update(self.phys, Uf, Uc, self.g, vp_, eps, dl, th)
        if lev + 1 < self.g.nt:
            if mistake == "wavelet_index":
                Uf[self.sz, self.sx] += self.src[lev]
            elif mistake == "source_sign":
                Uf[self.sz, self.sx] -= self.src[lev + 1]
        self.hi, self.lo = Uc, Uf
        self.lev = lev - 1
        return None

    monkeypatch.setattr(O._RandomRebuild, "_step", broken_step)
    rb = O._RandomRebuild(orc, 0, g, model, S, fw["current"], orc.arr(pb["src"]), sz, sx)
    scale = np.abs(S).max()
    worst = 0.0
    for t in range(g.nt - 2, 0, -1):
        St, _ = rb.get(t, False)
        worst = max(worst, float(np.abs(St - S[t]).max()))
    assert worst > 1e-3 * scale
