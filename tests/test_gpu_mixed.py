"""GPU: the mixed-precision path (complex FP32 LU on MFMA + FP64 iterative
refinement against the FP64 matrix) and the N = 16,020 grid.

Bar: Z_in within 1e-10 relative of the FP64 oracle (the north-star
tolerance for the mixed path), currents within 1e-9 in norm, and the
refinement's relative residual at most the 1e-12 it is asked for.
"""
import json
import os

import numpy as np
import pytest

from necpp_amd import decks
from necpp_amd import necpp as api
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

TOL_ZIN = 1e-10
TOL_CUR = 1e-9
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def rel(a, b):
    return abs(a - b) / abs(b)


def zin(ctx, k=0):
    return complex(api.nec_impedance_real(ctx, k), api.nec_impedance_imag(ctx, k))


def run(deck, precision, with_rp=True, tol=None):
    ctx = api.nec_create()
    for w in deck["wires"]:
        assert api.nec_wire(ctx, *w) == 0
    assert api.nec_geometry_complete(ctx, deck["gpflag"]) == 0, api.nec_error_message()
    assert api.nec_gpu_set_precision(ctx, precision) == 0
    if tol is not None:
        assert api.nec_gpu_set_refine_tolerance(ctx, tol) == 0
    assert api.nec_fr_card(ctx, 0, 1, deck["freq_mhz"], 0.0) == 0
    s = deck["source"]
    v = deck["voltage"]
    assert api.nec_ex_card(ctx, 0, s[0], s[1], 0, v.real, v.imag, 0, 0, 0, 0) == 0
    if with_rp and "rp" in deck:
        rc = api.nec_rp_card(ctx, *deck["rp"])
    else:
        rc = api.nec_xq_card(ctx, 0)
    return ctx, rc


def test_mixed_dipole_pattern_matches_oracle():
    deck = decks.dipole(51, rp=(37, 13))
    ctx, rc = run(deck, 1)
    try:
        assert rc == 0, api.nec_error_message()
        o = po.run_deck(deck, nthreads=4)
        assert rel(zin(ctx), o["zin"][0]) < TOL_ZIN
        cur = api.gpu_currents(ctx, 0)
        assert np.linalg.norm(cur - o["current"]) <= TOL_CUR * np.linalg.norm(o["current"])
        g = api.gpu_pattern_array(ctx, 0, "gain", 37 * 13)
        finite = o["pattern"]["gain"] > -999
        assert np.max(np.abs(g[finite] - o["pattern"]["gain"][finite])) < 1e-8
        steps, res = api.gpu_refinement(ctx)
        print("dipole mixed: refinement steps", steps, "residual", res)
        assert steps >= 1 and res <= 1e-12
    finally:
        api.nec_delete(ctx)


def test_mixed_c2_grid_matches_oracle():
    deck = decks.grid(41)
    ctx, rc = run(deck, 1, with_rp=False)
    try:
        assert rc == 0, api.nec_error_message()
        o = po.run_deck(deck, nthreads=16, with_rp=False)
        z = zin(ctx)
        print("C2 mixed Z_in", z, "oracle", o["zin"][0], "rel", rel(z, o["zin"][0]), api.gpu_refinement(ctx))
        assert rel(z, o["zin"][0]) < TOL_ZIN
        cur = api.gpu_currents(ctx, 0)
        assert np.linalg.norm(cur - o["current"]) <= TOL_CUR * np.linalg.norm(o["current"])
        steps, res = api.gpu_refinement(ctx)
        assert steps >= 1 and res <= 1e-12
    finally:
        api.nec_delete(ctx)


@pytest.mark.parametrize("mode", ["1", "0"])
def test_mixed_pivot_modes_match_oracle(mode, monkeypatch):
    """The FP32 factors with block-local pivots (the default) and with
    partial pivoting (NECGPU_MIXED_PIVOT=0) both refine to the oracle's
    Z_in; the grid's partial pivoting moves rows across 64-row blocks, so the
    block-local factors differ from the reference's and only the refined
    solution is compared."""
    monkeypatch.setenv("NECGPU_MIXED_PIVOT", mode)
    deck = decks.grid(21)
    ctx, rc = run(deck, 1, with_rp=False)
    try:
        assert rc == 0, api.nec_error_message()
        o = po.run_deck(deck, nthreads=8, with_rp=False)
        assert rel(zin(ctx), o["zin"][0]) < TOL_ZIN
        cur = api.gpu_currents(ctx, 0)
        assert np.linalg.norm(cur - o["current"]) <= TOL_CUR * np.linalg.norm(o["current"])
        steps, res = api.gpu_refinement(ctx)
        assert res <= 1e-12 and 1 <= steps <= 10, (steps, res)
    finally:
        api.nec_delete(ctx)


def test_precision_switch_refactors():
    # FP64 -> mixed -> FP64 on one context: every solve matches the oracle
    deck = decks.dipole(31)
    o = po.run_deck(deck, nthreads=2, with_rp=False)
    ctx, rc = run(deck, 0, with_rp=False)
    try:
        assert rc == 0
        assert rel(zin(ctx, 0), o["zin"][0]) < 1e-12
        assert api.nec_gpu_set_precision(ctx, 1) == 0
        assert api.nec_xq_card(ctx, 0) == 0, api.nec_error_message()
        assert rel(zin(ctx, 1), o["zin"][0]) < TOL_ZIN
        assert api.nec_gpu_set_precision(ctx, 0) == 0
        assert api.nec_xq_card(ctx, 0) == 0
        assert rel(zin(ctx, 2), o["zin"][0]) < 1e-12
        assert api.nec_gpu_set_precision(ctx, 2) == 1
        assert "precision" in api.nec_error_message()
    finally:
        api.nec_delete(ctx)


def test_refinement_limit_is_reported():
    # an unreachable tolerance: the solve stops after its step limit and the
    # card returns 1 with the residual it reached
    deck = decks.dipole(21)
    ctx, rc = run(deck, 1, with_rp=False, tol=1e-300)
    try:
        assert rc == 1
        assert "refinement did not reach" in api.nec_error_message()
    finally:
        api.nec_delete(ctx)
