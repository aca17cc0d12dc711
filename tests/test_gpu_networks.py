"""Networks and transmission lines (NT / TL cards) on the GPU path against
the oracle's netwk restatement (pinned in test_oracle_networks.py): one-
and two-port networks on undriven and driven segments, straight and crossed
lines with end admittances, line lengths from the port centres, a frequency
sweep, the pattern normalised by the power the networks take, and the
reference's errors."""
import numpy as np
import pytest
import torch  # noqa: F401  (imported at collection: the first import on a fresh box is slow)

from necpp_amd import necpp as api
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

TOL = 1e-10
WIRES = [(1, 11, 0.0, 0.0, -0.24, 0.0, 0.0, 0.24, 0.002, 1.0, 1.0),
         (2, 11, 0.15, 0.0, -0.23, 0.15, 0.0, 0.23, 0.002, 1.0, 1.0)]
# (card, tag1, seg1, tag2, seg2, six numbers); segment 6 of tag 1 is driven
CASES = {
    "one_port_undriven": [("NT", 2, 6, 2, 6, (0.01, -0.02, 0.0, 0.0, 0.0, 0.0))],
    "one_port_driven": [("NT", 1, 6, 1, 6, (0.004, 0.011, 0.0, 0.0, 0.0, 0.0))],
    "line": [("TL", 1, 6, 2, 6, (75.0, 0.37, 0.001, -0.002, 0.0, 0.003))],
    "crossed_line_from_centres": [("TL", 1, 3, 2, 9, (-50.0, 0.0, 0.0, 0.0, 0.0, 0.0))],
    "chain": [("NT", 1, 3, 2, 3, (0.002, 0.004, -0.001, 0.0025, 0.0015, -0.003)),
              ("TL", 2, 3, 2, 9, (60.0, 0.21, 0.0, 0.0, 0.0, 0.0)),
              ("TL", 1, 6, 1, 9, (300.0, 0.5, 0.0, 0.0, 0.0, -0.001))],
}


def build(freq=299.792458, nfreq=1, dfreq=0.0):
    ctx = api.nec_create()
    for w in WIRES:
        assert api.nec_wire(ctx, *w) == 0
    assert api.nec_geometry_complete(ctx, 0) == 0
    assert api.nec_fr_card(ctx, 0, nfreq, freq, dfreq) == 0
    assert api.nec_ex_card(ctx, 0, 1, 6, 0, 1.0, 0.0, 0, 0, 0, 0) == 0
    return ctx


def add_networks(ctx, nets):
    for card, t1, s1, t2, s2, x in nets:
        f = api.nec_tl_card if card == "TL" else api.nec_nt_card
        assert f(ctx, t1, s1, t2, s2, *x) == 0, api.nec_error_message()


def oracle(nets, f, rp=None):
    g = po.Geometry()
    for w in WIRES:
        g.wire(*w)
    g.complete(0)
    g.set_networks([(c, g.segment_number(t1, s1), g.segment_number(t2, s2), x) for c, t1, s1, t2, s2, x in nets])
    return g.run_voltage(f, [(6, 1.0)], rp=rp, nthreads=8)


@pytest.mark.parametrize("case", list(CASES))
def test_networks_match_oracle(case):
    nets = CASES[case]
    f = 299.792458
    # RP with ipd = 1: the gain normalised by the radiated power, the input
    # power less what the networks take
    rp = dict(n_theta=19, n_phi=4, theta0=0.0, phi0=0.0, dtheta=10.0, dphi=90.0, output_format=1, normalization=0,
              ipd=1, iavp=0)
    ctx = build(f)
    try:
        add_networks(ctx, nets)
        assert api.nec_rp_card(ctx, 0, 19, 4, 1, 0, 1, 0, 0.0, 0.0, 10.0, 90.0, 0.0, 0.0) == 0, api.nec_error_message()
        o = oracle(nets, f, rp)
        z = complex(api.nec_impedance_real(ctx, 0), api.nec_impedance_imag(ctx, 0))
        assert abs(z - o["zin"][0]) <= TOL * abs(o["zin"][0]), (z, o["zin"][0])
        cur = api.gpu_currents(ctx, 0)
        assert np.linalg.norm(cur - o["current"]) <= TOL * np.linalg.norm(o["current"])
        gain = api.gpu_pattern_array(ctx, 0, "gain", 76)
        finite = o["pattern"]["gain"] > -999.0
        assert np.array_equal(gain > -999.0, finite)
        assert np.max(np.abs(gain[finite] - o["pattern"]["gain"][finite])) < 1e-8
    finally:
        api.nec_delete(ctx)


def test_networks_over_a_frequency_sweep():
    # a line of length 0 takes the distance of the port centres at the first
    # frequency and keeps it
    nets = CASES["crossed_line_from_centres"] + CASES["one_port_undriven"]
    freqs = (250.0, 280.0, 310.0)
    ctx = build(freqs[0], 3, 30.0)
    try:
        add_networks(ctx, nets)
        assert api.nec_xq_card(ctx, 0) == 0, api.nec_error_message()
        g = po.Geometry()
        for w in WIRES:
            g.wire(*w)
        g.complete(0)
        g.set_networks([(c, g.segment_number(t1, s1), g.segment_number(t2, s2), x) for c, t1, s1, t2, s2, x in nets])
        for k, f in enumerate(freqs):
            o = g.run_voltage(f, [(6, 1.0)], nthreads=8)
            z = complex(api.nec_impedance_real(ctx, k), api.nec_impedance_imag(ctx, k))
            assert abs(z - o["zin"][0]) <= TOL * abs(o["zin"][0]), (f, z, o["zin"][0])
    finally:
        api.nec_delete(ctx)


def test_network_cards_replace_and_clear():
    f = 299.792458
    ctx = build(f)
    try:
        # a second run of TL cards after another card replaces the first
        add_networks(ctx, CASES["line"])
        assert api.nec_xq_card(ctx, 0) == 0
        add_networks(ctx, CASES["one_port_undriven"])
        assert api.nec_xq_card(ctx, 0) == 0
        o = oracle(CASES["one_port_undriven"], f)
        z = complex(api.nec_impedance_real(ctx, 1), api.nec_impedance_imag(ctx, 1))
        assert abs(z - o["zin"][0]) <= TOL * abs(o["zin"][0])
        # a port segment of -1 clears them
        assert api.nec_tl_card(ctx, 0, -1, 0, 0, 0, 0, 0, 0, 0, 0) == 0
        assert api.nec_xq_card(ctx, 0) == 0
        o = oracle([], f)
        z = complex(api.nec_impedance_real(ctx, 2), api.nec_impedance_imag(ctx, 2))
        assert abs(z - o["zin"][0]) <= TOL * abs(o["zin"][0])
    finally:
        api.nec_delete(ctx)


def test_source_missing_from_the_network_table_is_an_error():
    ctx = build()
    try:
        assert api.nec_ex_card(ctx, 0, 2, 6, 0, 1.0, 0.0, 0, 0, 0, 0) == 0  # a second source
        add_networks(ctx, CASES["one_port_driven"])
        assert api.nec_xq_card(ctx, 0) == 1
        assert "NETWORK DATA ERROR" in api.nec_error_message()
    finally:
        api.nec_delete(ctx)
