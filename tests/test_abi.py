"""The C-ABI library: symbols, host-side card handling and error behaviour.

Runs without a GPU: nec_create and the geometry cards are host code; the
device is first touched by XQ / RP.  Geometry tables are compared with the
oracle bit for bit (junction averaging, centres, lengths, direction cosines).
"""
import os
import re

import numpy as np
import pytest

from necpp_amd import decks
from necpp_amd import necpp as api
from oracle import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in ("libnecpp.h", "necpp_gpu.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(nec_[a-z0-9_]+|necgpu_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    L = api.lib()
    declared = declared_functions()
    assert "nec_create" in declared and "nec_rp_card" in declared
    missing = [n for n in sorted(declared) if not hasattr(L, n)]
    assert not missing, missing
    # and the binding knows every declared entry point
    assert declared <= set(api.EXPORTS)


def reference_libnecpp_names():
    # the reference header's function names (src/libnecpp.h), listed here so
    # the test runs where the reference tree is absent
    return ["nec_create", "nec_delete", "nec_wire", "nec_sp_card", "nec_sc_card", "nec_gm_card", "nec_gx_card",
            "nec_geometry_complete", "nec_benchmark", "nec_error_message", "nec_medium_parameters", "nec_gn_card",
            "nec_fr_card", "nec_ek_card", "nec_ld_card", "nec_ex_card", "nec_excitation_voltage",
            "nec_excitation_current", "nec_excitation_planewave", "nec_tl_card", "nec_nt_card", "nec_xq_card",
            "nec_gd_card", "nec_rp_card", "nec_pt_card", "nec_pq_card", "nec_kh_card", "nec_ne_card", "nec_nh_card",
            "nec_cp_card", "nec_pl_card", "nec_gain", "nec_gain_max", "nec_gain_min", "nec_gain_mean",
            "nec_gain_sd", "nec_gain_rhcp_max", "nec_gain_rhcp_min", "nec_gain_rhcp_mean", "nec_gain_rhcp_sd",
            "nec_gain_lhcp_max", "nec_gain_lhcp_min", "nec_gain_lhcp_mean", "nec_gain_lhcp_sd",
            "nec_impedance_real", "nec_impedance_imag", "nec_impedance_real_indexed",
            "nec_impedance_imag_indexed"]


def test_drop_in_covers_reference_api():
    L = api.lib()
    for n in reference_libnecpp_names():
        assert hasattr(L, n), n


def make(deck):
    ctx = api.nec_create()
    assert ctx
    for w in deck["wires"]:
        assert api.nec_wire(ctx, *w) == 0
    assert api.nec_geometry_complete(ctx, deck["gpflag"]) == 0, api.nec_error_message()
    return ctx


@pytest.mark.parametrize("deck", [decks.dipole(7, rp=None), decks.dipole(51, rp=None), decks.grid(9, rp=None),
                                  decks.yagi([0.5, 0.48, 0.45, 0.44], [0.4, 0.3, 0.35], 0.003, 11, rp=None)],
                         ids=["dipole7", "dipole51", "grid9", "yagi4"])
def test_geometry_tables_match_oracle_bitwise(deck):
    ctx = make(deck)
    try:
        tab, ic1, ic2, tags = api.gpu_segments(ctx)
        g = po.geometry_from_deck(deck)
        otab, oic1, oic2, otags = g.segments()
        assert np.array_equal(tab, otab)
        assert np.array_equal(ic1, oic1) and np.array_equal(ic2, oic2) and np.array_equal(tags, otags)
    finally:
        api.nec_delete(ctx)


def test_grid_junctions_connect_four_ends():
    ctx = make(decks.grid(5, rp=None))
    try:
        tab, ic1, ic2, _ = api.gpu_segments(ctx)
        n = tab.shape[0]
        assert n == 2 * 5 * 4
        # every interior edge end is connected; corner nodes have two ends
        assert np.count_nonzero(ic1 == 0) + np.count_nonzero(ic2 == 0) == 0
    finally:
        api.nec_delete(ctx)


def test_error_conventions():
    ctx = api.nec_create()
    try:
        # empty geometry
        assert api.nec_geometry_complete(ctx, 0) == 1
        assert "no wires" in api.nec_error_message()
        assert api.nec_wire(ctx, 1, 5, 0, 0, -1, 0, 0, 1, 0.01, 1, 1) == 0
        assert api.nec_geometry_complete(ctx, 0) == 0
        # unknown tag, segment out of range
        assert api.nec_ex_card(ctx, 0, 7, 1, 0, 1.0, 0, 0, 0, 0, 0) == 1
        assert "ITAG" in api.nec_error_message()
        assert api.nec_ex_card(ctx, 0, 0, 9, 0, 1.0, 0, 0, 0, 0, 0) == 1
        assert "TOO LARGE" in api.nec_error_message()
        # cards outside the GPU path fail loudly
        assert api.nec_cp_card(ctx, 1, 1, 1, 5) == 1
        assert "not supported" in api.nec_error_message()
        # TL on the path: a port segment that does not exist is a data fault
        assert api.nec_tl_card(ctx, 1, 9, 1, 5, 50.0, 0, 0, 0, 0, 0) == 1
        # LD runs on the GPU path; a bad tag range is a data fault
        assert api.nec_ld_card(ctx, 0, 0, 1, 1, 10.0, 0, 0) == 0
        assert api.nec_ld_card(ctx, 0, 0, 3, 2, 10.0, 0, 0) == 1
        assert "ITAG STEP1" in api.nec_error_message()
        assert api.nec_gn_card(ctx, 2, 0, 10.0, 0.01, 0, 0, 0, 0) == 1  # Sommerfeld ground
        assert "not supported" in api.nec_error_message()
        assert api.nec_gn_card(ctx, 1, 0, 0, 0, 0, 0, 0, 0) == 0  # perfect ground
        assert api.nec_gn_card(ctx, -1, 0, 0, 0, 0, 0, 0, 0) == 0
        assert api.nec_rp_card(ctx, 1, 10, 1, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0) == 1
        # missing results give the -999 sentinel
        assert api.nec_impedance_real(ctx, 0) == -999.0
        assert api.nec_impedance_imag_indexed(ctx, 0, 3) == -999.0
        assert api.nec_gain(ctx, 0, 0, 0) == -999.0
        assert api.nec_gain_max(ctx, 2) == -999.0
    finally:
        assert api.nec_delete(ctx) == 0


def test_segment_data_error():
    ctx = api.nec_create()
    try:
        assert api.nec_wire(ctx, 1, 4, 0, 0, 0, 0, 0, 1, 0.0, 1, 1) == 0
        assert api.nec_geometry_complete(ctx, 0) == 1
        assert "SEGMENT DATA ERROR" in api.nec_error_message()
    finally:
        api.nec_delete(ctx)


def test_ground_touching_wire_rules():
    # gpflag = 1: an end below ground is an error (build_connections)
    ctx = api.nec_create()
    try:
        assert api.nec_wire(ctx, 1, 4, 0, 0, -0.5, 0, 0, 1, 0.001, 1, 1) == 0
        assert api.nec_geometry_complete(ctx, 1) == 1
        assert "BELOW GROUND" in api.nec_error_message()
    finally:
        api.nec_delete(ctx)
