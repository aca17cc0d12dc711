"""GPU parity of every far-field pattern output on an elliptically polarised
structure: the twelve pattern arrays (gain, vertical/major, horizontal/minor,
total, RHCP, LHCP, axial ratio, tilt, sense, averaging scale, E_theta,
E_phi), the twelve statistics nec_gain{,_rhcp,_lhcp}_{max,min,mean,sd}
(src/libnecpp.h:547-589, src/nec_radiation_pattern.h:252-333) and the power
average (RP's A = 1), against the oracle's pattern (analyze,
src/nec_radiation_pattern.cpp:209-450), whose polarisation outputs are
pinned by tests/test_oracle_polarisation.py.

The structure is a turnstile: two crossed half-wave dipoles fed in
quadrature, so the polarisation is elliptical in most directions and RHCP
and LHCP differ by up to tens of dB (a z-directed dipole would leave them
equal and a swapped split unseen).
"""
import numpy as np
import pytest
import torch  # noqa: F401  (imported at collection: the first import on a fresh box is slow)

from gpu_compare import STATS, check_stats, compare_pattern  # noqa: F401
from necpp_amd import decks
from necpp_amd import necpp as api
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

def gpu_turnstile(phase, rp_card, precision=0):
    deck = decks.turnstile(phase)
    ctx = api.nec_create()
    for w in deck["wires"]:
        assert api.nec_wire(ctx, *w) == 0
    assert api.nec_geometry_complete(ctx, 0) == 0
    assert api.nec_gpu_set_precision(ctx, precision) == 0
    assert api.nec_fr_card(ctx, 0, 1, deck["freq_mhz"], 0.0) == 0
    for (tag, seg), v in deck["sources"]:
        assert api.nec_ex_card(ctx, 0, tag, seg, 0, v.real, v.imag, 0, 0, 0, 0) == 0
    assert api.nec_rp_card(ctx, *rp_card) == 0, api.nec_error_message()
    return ctx, deck


@pytest.mark.parametrize("fmt", [0, 1])
@pytest.mark.parametrize("phase", [-1j, 1j])
def test_turnstile_polarisation_matches_oracle(fmt, phase):
    # RP 0 19 37 <fmt>0 1 0 0 5 10: upper hemisphere (where the two senses
    # dominate differently), power average
    rp_card = (0, 19, 37, fmt, 0, 0, 1, 0.0, 0.0, 5.0, 10.0, 0.0, 0.0)
    ctx, deck = gpu_turnstile(phase, rp_card)
    try:
        g = po.geometry_from_deck(deck)
        src = [(g.segment_number(*ts), v) for ts, v in deck["sources"]]
        o = g.run_voltage(deck["freq_mhz"], src, rp=po.rp_from_card(rp_card), nthreads=4)
        for f in range(2):
            z = complex(api.nec_impedance_real_indexed(ctx, 0, f), api.nec_impedance_imag_indexed(ctx, 0, f))
            assert abs(z - o["zin"][f]) < 1e-10 * abs(o["zin"][f])
        gp = compare_pattern(ctx, o, 19 * 37)
        # elliptical polarisation almost everywhere: the comparison means something
        assert np.mean(gp["axial_ratio"] > 0.05) > 0.5
        assert np.max(np.abs(gp["pg_rhcp"] - gp["pg_lhcp"])) > 10.0
        got = check_stats(ctx, o, gp)
        # the RHCP and LHCP statistics really differ here (a swap would show)
        assert abs(got[4] - got[8]) > 3.0 and abs(got[6] - got[10]) > 3.0
        avg = np.zeros(1)
        sa = np.zeros(1)
        assert api.nec_gpu_pattern_average(ctx, 0, api._ptr(avg), api._ptr(sa)) == 0
        assert abs(avg[0] - o["average_power_gain"]) < 1e-10 * abs(o["average_power_gain"])
    finally:
        api.nec_delete(ctx)


def test_turnstile_polarisation_mixed_precision():
    # the mixed path's pattern from its refined currents, same bar
    rp_card = (0, 19, 37, 0, 0, 0, 0, 0.0, 0.0, 10.0, 10.0, 0.0, 0.0)
    ctx, deck = gpu_turnstile(-1j, rp_card, precision=1)
    try:
        g = po.geometry_from_deck(deck)
        src = [(g.segment_number(*ts), v) for ts, v in deck["sources"]]
        o = g.run_voltage(deck["freq_mhz"], src, rp=po.rp_from_card(rp_card), nthreads=4)
        check_stats(ctx, o, compare_pattern(ctx, o, 19 * 37))
    finally:
        api.nec_delete(ctx)


# ------------------------------------------------ patterns at config size

def run_grid(nn, precision, rp=(91, 361)):
    deck = decks.grid(nn, rp=rp)
    ctx = api.nec_create()
    for w in deck["wires"]:
        assert api.nec_wire(ctx, *w) == 0
    assert api.nec_geometry_complete(ctx, 0) == 0
    assert api.nec_gpu_set_precision(ctx, precision) == 0
    assert api.nec_fr_card(ctx, 0, 1, deck["freq_mhz"], 0.0) == 0
    s = deck["source"]
    assert api.nec_ex_card(ctx, 0, s[0], s[1], 0, 1.0, 0.0, 0, 0, 0, 0) == 0
    rc = api.nec_rp_card(ctx, *deck["rp"])
    return ctx, deck, rc


def test_c2_grid_full_pattern_matches_oracle():
    # N = 3,280, the bench's workload: all twelve 91 x 361 arrays and the
    # statistics against the oracle's own solve and pattern
    ctx, deck, rc = run_grid(41, 0)
    try:
        assert rc == 0, api.nec_error_message()
        o = po.run_deck(deck, nthreads=16)
        z = complex(api.nec_impedance_real(ctx, 0), api.nec_impedance_imag(ctx, 0))
        assert abs(z - o["zin"][0]) < 1e-10 * abs(o["zin"][0])
        check_stats(ctx, o, compare_pattern(ctx, o, 91 * 361))
        assert abs(api.nec_gain_max(ctx, 0) - o["max_gain"]) < 1e-9
    finally:
        api.nec_delete(ctx)


def test_c5_grid_pattern_fp64_and_mixed_match_oracle_fixture():
    # N = 16,020 (4.1 GB FP64 matrix), FP64 and mixed precision: Z_in,
    # currents and the 91 x 361 pattern against the oracle's solve, whose
    # current coefficients are committed (make_c5_oracle.py); the oracle
    # pattern is recomputed here from them
    import json
    import os
    golden = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    f = json.load(open(os.path.join(golden, "c5_oracle.json")))
    ocur = np.load(os.path.join(golden, "c5_oracle_current.npy"))
    abc = np.load(os.path.join(golden, "c5_oracle_abc.npy"))
    opat = np.load(os.path.join(golden, "c5_oracle_pattern.npy"))
    deck = decks.grid(90)
    g = po.geometry_from_deck(deck)
    o = g.pattern(abc, deck["freq_mhz"], f["input_power"], po.rp_from_card(deck["rp"]), nthreads=16)
    # the fixture's pattern is the one the oracle computed in its own solve
    assert np.array_equal(o["pattern"]["pg_tot"], opat)
    assert np.allclose(o["stats"], f["pattern_stats"], rtol=0, atol=0)
    for precision in (0, 1):
        ctx, _, rc = run_grid(90, precision)
        try:
            assert rc == 0, api.nec_error_message()
            z = complex(api.nec_impedance_real(ctx, 0), api.nec_impedance_imag(ctx, 0))
            zo = complex(*f["zin"])
            cur = api.gpu_currents(ctx, 0)
            cur_err = np.linalg.norm(cur - ocur) / np.linalg.norm(ocur)
            gp = {k: api.gpu_pattern_array(ctx, 0, k, 91 * 361) for k in ("pg_tot", "e_theta", "e_phi")}
            print(f"C5 precision {precision}: Z_in rel err {abs(z - zo) / abs(zo):.2e}, currents {cur_err:.2e}, "
                  f"total gain max abs dB err {np.max(np.abs(gp['pg_tot'] - opat)):.2e}, E_theta rel "
                  f"{np.max(np.abs(gp['e_theta'] - o['pattern']['e_theta'])) / np.max(np.abs(o['pattern']['e_theta'])):.2e}",
                  flush=True)
            assert abs(z - zo) < 1e-10 * abs(zo)
            assert cur_err <= (1e-12 if precision == 0 else 1e-9)
            check_stats(ctx, o, compare_pattern(ctx, o, 91 * 361))
        finally:
            api.nec_delete(ctx)
