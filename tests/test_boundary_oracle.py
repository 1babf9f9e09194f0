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
