"""Pins of the oracle's restatements of the library's own storage strategies
(no counterpart in the reference): binomial checkpointing driven step by step
in the library's convention reproduces save-every-step bit for bit with
float32 slots, for any number of slots, and never advances a step more often
than the treeverse bound allows; its float16 slots and the 16-bit save-all
snapshots stay within the rounding their formats allow."""
from math import comb

import numpy as np
import pytest

from oracle import oracle as O
from oracle.oracle import Oracle, rel_l2
from tests.problems import make_problem


def _setup(phys, abc, nt=90, seed=5):
    pb = make_problem(phys, abc, nz=24, nx=30, N=10, nt=nt, itlag=10, seed=seed)
    orc = Oracle("f32")
    g = pb["g"]
    model = tuple(orc.expand(m, g.N) if m is not None else None for m in (pb["vp"], pb["eps"], pb["dl"], pb["th"]))
    A = orc.arr(orc.cerjan(g.N)) if abc == 0 else orc.zeros(max(g.N, 1))
    cpml = (pb["d0"], pb["fp"]) if abc == 1 else None
    fw = orc.forward_shot(phys, abc, g, model, pb["src"], int(pb["sz"][0]), int(pb["sx"][0]), pb["rz"], pb["rx"],
                          A=A, cpml=cpml, save=True)
    return pb, orc, g, model, A, cpml, fw["save"]


@pytest.mark.parametrize("phys,abc", [(0, 0), (0, 1), (1, 0), (1, 1), (2, 1)])
@pytest.mark.parametrize("slots", [1, 4, 50])
def test_revolve_float32_is_save_all_bit_for_bit(phys, abc, slots):
    pb, orc, g, model, A, cpml, S = _setup(phys, abc)
    t_lo, t_hi = 1, g.nt - 2
    steps = [0]
    orig = O._Revolve._step

    def counting(self, k, save=False):
        steps[0] += 1
        return orig(self, k, save)

    O._Revolve._step = counting
    try:
        rv = O._Revolve(orc, phys, abc, g, model, A, cpml, orc.arr(pb["src"]), int(pb["sz"][0]),
                        int(pb["sx"][0]), t_lo, t_hi, slots, "f32")
    finally:
        O._Revolve._step = orig
    sa = O._SaveAll(S, g)
    for t in range(t_lo, t_hi + 1):
        s_rv, d_rv = rv.get(t, True)
        s_sa, d_sa = sa.get(t, True)
        assert np.array_equal(s_rv, s_sa), t
        assert np.array_equal(d_rv, d_sa), t
    n = t_hi - t_lo + 1
    s = max(1, min(slots, n)) - 1
    r = n  # no free slot: every adjoint step recomputed from the range's start
    if s > 0:
        r = 0
        while comb(s + r, s) < n:
            r += 1
    # the first descent, then at most r recomputations per adjoint step, plus the leaf steps
    assert steps[0] <= t_lo + r * n + n


def test_revolve_split_respects_the_binomial_bound():
    for s in range(1, 8):
        for n in range(2, 300):
            m = O.revolve_split(n, s)
            r = 1
            while comb(s + r, s) < n:
                r += 1
            assert 1 <= m < n
            assert m <= comb(s + r - 1, s) and n - m <= comb(s - 1 + r, s - 1)


@pytest.mark.parametrize("abc", [0, 1])
def test_revolve_float16_slots_stay_near_float32(abc):
    pb, orc, g, model, A, cpml, S = _setup(0, abc, nt=120)
    args = (orc, 0, abc, g, model, A, cpml, orc.arr(pb["src"]), int(pb["sz"][0]), int(pb["sx"][0]), 1, g.nt - 2, 6)
    r16, r32 = O._Revolve(*args, "f16"), O._Revolve(*args, "f32")
    sl = (slice(g.N, g.N + g.nz), slice(g.N, g.N + g.nx))
    es = max(rel_l2(r16.get(t, True)[0][sl], r32.get(t, True)[0][sl]) for t in range(1, g.nt - 1))
    ed = rel_l2(np.stack([r16.get(t, True)[1] for t in range(1, g.nt - 1)]),
                np.stack([r32.get(t, True)[1] for t in range(1, g.nt - 1)]))
    # S_t: a few float16 roundings (2^-11 of each field's max) carried forward
    assert 0 < es < 5e-3
    # D_t: the operator term of the restored state amplifies its rounding, but stays bounded
    assert 0 < ed < 1e-1


def _half_ulp_at(q, kind):
    """Half an ulp of float16 (10 fraction bits) / bfloat16 (7) at |q|."""
    p = 10 if kind == "f16" else 7
    return 0.5 * 2.0 ** (np.floor(np.log2(q)) - p)


@pytest.mark.parametrize("kind", ["f16", "bf16"])
def test_half_snapshots_round_within_their_format(kind):
    pb, orc, g, model, A, cpml, S = _setup(1, 1, nt=80)
    sa, sh = O._SaveAll(S, g), O._SaveAll(S, g, kind)
    sl = (slice(g.N, g.N + g.nz), slice(g.N, g.N + g.nx))
    for t in range(2, g.nt - 2, 7):
        s32 = sa.get(t, False)[0][sl]
        s16 = sh.get(t, False)[0][sl]
        m = np.max(np.abs(s32))
        # each sample within half an ulp of the format at the snapshot's scale (max |S_t|: |q| <= 1)
        assert np.max(np.abs(s16 - s32)) <= _half_ulp_at(1.0, kind) * m * (1 + 1e-6)
        assert np.array_equal(sh.get(t, False)[0][:g.N], sa.get(t, False)[0][:g.N])  # absorbing layer untouched
        d32, d16 = sa.get(t, True)[1], sh.get(t, True)[1]
        m = np.max(np.abs(sa.get(t - 1, True)[1]))  # the lagged scale: |q| may exceed 1
        qmax = np.max(np.abs(d32)) / m
        assert np.max(np.abs(d16 - d32)) <= _half_ulp_at(qmax, kind) * m * (1 + 1e-6)
        assert 0 < np.max(np.abs(d16 - d32))


def test_a_scale_below_the_smallest_normal_float_stores_zeros():
    """A field whose max |v| is a subnormal float (CPML memory late in a long
    run): 1 / max would overflow to Inf and 0 * Inf is NaN, so such a field is
    stored as zeros (the library's scale_inv does the same)."""
    from oracle import exact_adjoint as EA
    from oracle.oracle import quantize_scaled

    v = np.array([1e-39, -5e-40, 0.0, 2e-45], dtype=np.float32)
    for q in (quantize_scaled(v, "f16"), quantize_scaled(v, "bf16"), EA._quantize(v, "f16")):
        assert np.all(np.isfinite(q)) and not np.any(q)
    w = np.array([2e-38, 1e-38, 0.0], dtype=np.float32)  # smallest normals still scale
    assert np.isfinite(quantize_scaled(w, "f16")).all() and quantize_scaled(w, "f16")[0] == np.float32(2e-38)
