"""The card-deck front end (nec_gpu_deck) without a GPU: line and field
rules, the geometry section (GW, GC, GS, GE) against the same wires through
the API, and the error paths.  Decks stop before XQ / RP, which would touch
the device."""
import numpy as np
import pytest

from necpp_amd import necpp as api


def run(text):
    ctx = api.nec_create()
    rc = api.nec_gpu_deck(ctx, text.encode())
    return ctx, rc


def segs(ctx):
    return api.gpu_segments(ctx)


def api_ctx(wires, gpflag=0):
    ctx = api.nec_create()
    for w in wires:
        assert api.nec_wire(ctx, *w) == 0
    assert api.nec_geometry_complete(ctx, gpflag) == 0
    return ctx


def same_geometry(a, b):
    ta, ia1, ia2, ga = segs(a)
    tb, ib1, ib2, gb = segs(b)
    return np.array_equal(ta, tb) and np.array_equal(ia1, ib1) and np.array_equal(ia2, ib2) and np.array_equal(ga, gb)


def test_geometry_section_matches_api():
    text = """CM two wires, one tapered
CM second comment line
CE
GW 1 7 0 0 -0.25 0 0 0.25 0.001
GW 2 5 0.1 0 -0.2 0.1 0 0.2 0
GC 0 0 1.1 0.001 0.002
GE 0
FR 0 1 0 0 299.8
EN
"""
    ctx, rc = run(text)
    ref = api_ctx([(1, 7, 0, 0, -0.25, 0, 0, 0.25, 0.001, 1.0, 1.0),
                   (2, 5, 0.1, 0, -0.2, 0.1, 0, 0.2, 0.001, 1.1, (0.002 / 0.001) ** (1 / 4.0))])
    try:
        assert rc == 0, api.nec_error_message()
        assert same_geometry(ctx, ref)
    finally:
        api.nec_delete(ctx)
        api.nec_delete(ref)


def test_gs_scales_wires_defined_before_it():
    text = "CE\nGW 1 5 0 0 -1 0 0 1 0.01\nGS 0 0 0.5\nGW 2 5 1 0 -1 1 0 1 0.01\nGE 0\nEN\n"
    ctx, rc = run(text)
    ref = api_ctx([(1, 5, 0, 0, -0.5, 0, 0, 0.5, 0.005, 1.0, 1.0), (2, 5, 1, 0, -1, 1, 0, 1, 0.01, 1.0, 1.0)])
    try:
        assert rc == 0, api.nec_error_message()
        assert same_geometry(ctx, ref)
    finally:
        api.nec_delete(ctx)
        api.nec_delete(ref)


def test_line_rules():
    # '#' lines and lines starting with a blank are skipped, a quote starts a
    # comment, commas separate fields, mnemonics are case-insensitive, no CM/CE
    text = "# a comment line\n gw this line is skipped\ngw 1,7, 0,0,-0.25, 0,0,0.25, 0.001 ' trailing\r\nge 0\r\nen\r\n"
    ctx, rc = run(text)
    ref = api_ctx([(1, 7, 0, 0, -0.25, 0, 0, 0.25, 0.001, 1.0, 1.0)])
    try:
        assert rc == 0, api.nec_error_message()
        assert same_geometry(ctx, ref)
    finally:
        api.nec_delete(ctx)
        api.nec_delete(ref)


@pytest.mark.parametrize("text,needle", [
    ("CM no end of comments\nGW 1 5 0 0 -1 0 0 1 0.01\nGE 0\n", "COMMENT CARD"),
    ("CE\nGW 1 5 0 0 -1 0 0 1 0.01\nGR 0 4\nGE 0\n", "GR: not supported"),
    ("CE\nGW 1 5 0 0 -1 0 0 1 0.01\nGZ 0 0\nGE 0\n", "GEOMETRY DATA CARD ERROR"),
    ("CE\nGW 1 5 0 0 -1 0 0 1 0.0x1\nGE 0\n", "NON-NUMERICAL CHARACTER 'x'"),
    ("CE\nGW 1 5 0 0 -1 0 0 1 0\nGE 0\n", "GEOMETRY DATA CARD ERROR"),
    ("CE\nGW 1 5 0 0 -1 0 0 1 0.01\n", "no GE card"),
    ("CE\nGW 1 5 0 0 -1 0 0 1 0.01\nGE 0\nQQ 1\n", "FAULTY DATA CARD LABEL"),
    ("CE\nGW 1 5 0 0 -1 0 0 1 0.01\nGE 0\nCP 1 3 1 3\n", "CP: not supported"),
    ("CE\nGW 1 5 0 0 -1 0 0 1 0.01\nGE 0\nTL 1 3 2 3 50\n", "ITAG"),
    ("CE\nGW 1 5 0 0 -1 0 0 1 0.01\nGE 0\nEX 0 7 1 0 1\n", "ITAG"),
    ("CE\nGW 1 5 0 0 -1 0 0 1 0.01\nGE 0\nGN 2 0 0 0 10 0.01\n", "GN"),
])
def test_deck_errors(text, needle):
    ctx, rc = run(text)
    try:
        assert rc == 1
        msg = api.nec_error_message()
        assert needle in msg, msg
    finally:
        api.nec_delete(ctx)


def test_error_names_the_deck_line():
    ctx, rc = run("CE\nGW 1 5 0 0 -1 0 0 1 0.01\nGE 0\nFR 0 1 0 0 300\nCP 1 3 1 3\n")
    try:
        assert rc == 1 and "deck line 5 (CP)" in api.nec_error_message()
    finally:
        api.nec_delete(ctx)


def oracle_geometry(cards, gpflag=0):
    """The oracle's geometry from a list of ("GW", wire) / ("GM", args) /
    ("GS", f) cards, applied in order to its segments."""
    from oracle import pyoracle as po
    g = po.Geometry()
    for kind, a in cards:
        if kind == "GW":
            g.wire(*a)
        elif kind == "GM":
            g.gm(*a)
        else:
            g.gs(a)
    g.complete(gpflag)
    return g


def same_as_oracle(ctx, g):
    tab, ic1, ic2, tags = segs(ctx)
    otab, oic1, oic2, otags = g.segments()
    return (np.array_equal(tab, otab) and np.array_equal(ic1, oic1) and np.array_equal(ic2, oic2)
            and np.array_equal(tags, otags))


def test_gm_copies_and_moves():
    # GM acts on segments (c_geometry::move): the deck, the API call and the
    # oracle's restatement give the same segment tables bit for bit
    text = """CE
GW 1 3 0 0 0 0 0 0.1 0.001
GW 2 3 0 0.1 0 0 0.1 0.1 0.001
GM 5 3 0 0 30 0.2 0 0 2
GS 0 0 0.9
GM 0 0 10 20 0 0 0 0.05 0
GE 0
EN
"""
    ctx, rc = run(text)
    w1 = (1, 3, 0, 0, 0, 0, 0, 0.1, 0.001, 1.0, 1.0)
    w2 = (2, 3, 0, 0.1, 0, 0, 0.1, 0.1, 0.001, 1.0, 1.0)
    g = oracle_geometry([("GW", w1), ("GW", w2), ("GM", (5, 3, 0, 0, 30, 0.2, 0, 0, 2)), ("GS", 0.9),
                         ("GM", (0, 0, 10, 20, 0, 0, 0, 0.05, 0))])
    try:
        assert rc == 0, api.nec_error_message()
        assert same_as_oracle(ctx, g)
        _, _, _, tags = segs(ctx)
        assert list(tags) == [1] * 3 + [2] * 3 + [7] * 3 + [12] * 3 + [17] * 3
    finally:
        api.nec_delete(ctx)


def test_gm_api_matches_deck():
    text = "CE\nGW 1 4 0 0 0 0.4 0 0 0.001\nGM 3 2 0 0 45 0 0 0.1 1\nGM 0 0 0 90 0 0.5 0 0 4\nGE 0\nEN\n"
    ctx, rc = run(text)
    ref = api.nec_create()
    try:
        assert rc == 0, api.nec_error_message()
        assert api.nec_wire(ref, 1, 4, 0, 0, 0, 0.4, 0, 0, 0.001, 1.0, 1.0) == 0
        assert api.nec_gm_card(ref, 3, 2, 0, 0, 45, 0, 0, 0.1, 1) == 0, api.nec_error_message()
        assert api.nec_gm_card(ref, 0, 0, 0, 90, 0, 0.5, 0, 0, 4) == 0, api.nec_error_message()
        assert api.nec_geometry_complete(ref, 0) == 0
        assert same_geometry(ctx, ref)
        _, _, _, tags = segs(ref)
        assert list(tags) == [1] * 4 + [4] * 4 + [7] * 4
    finally:
        api.nec_delete(ctx)
        api.nec_delete(ref)


def test_gm_tag_that_no_segment_has_is_an_error():
    # c_geometry::move looks the tag up with get_segment_number
    # (src/c_geometry.cpp:76-98, :865), which throws for an unknown tag
    ctx, rc = run("CE\nGW 1 4 0 0 0 0.4 0 0 0.001\nGM 0 1 0 0 45 0 0 0 9\nGE 0\nEN\n")
    try:
        assert rc == 1
        msg = api.nec_error_message()
        assert "NO SEGMENT HAS AN ITAG OF 9" in msg and "deck line 3 (GM)" in msg, msg
    finally:
        api.nec_delete(ctx)
    ctx = api.nec_create()
    try:
        assert api.nec_wire(ctx, 1, 4, 0, 0, 0, 0.4, 0, 0, 0.001, 1.0, 1.0) == 0
        assert api.nec_gm_card(ctx, 0, 1, 0, 0, 45, 0, 0, 0, 9) == 1
        assert "NO SEGMENT HAS AN ITAG OF 9" in api.nec_error_message()
    finally:
        api.nec_delete(ctx)


