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
    ("CE\nGW 1 5 0 0 -1 0 0 1 0.01\nGE 0\nTL 1 3 1 3 50\n", "TL: not supported"),
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
    ctx, rc = run("CE\nGW 1 5 0 0 -1 0 0 1 0.01\nGE 0\nFR 0 1 0 0 300\nTL 1 3 1 3 50\n")
    try:
        assert rc == 1 and "deck line 5 (TL)" in api.nec_error_message()
    finally:
        api.nec_delete(ctx)


KPI = 3.1415926536


def gm(wires, itgi, nrpt, rox, roy, roz, xs, ys, zs, its=0):
    """GM on wire endpoints, as the deck front end applies it."""
    import math
    d2r = lambda d: d * ((2 * KPI) / 360.0)
    sps, cps = math.sin(d2r(rox)), math.cos(d2r(rox))
    sth, cth = math.sin(d2r(roy)), math.cos(d2r(roy))
    sph, cph = math.sin(d2r(roz)), math.cos(d2r(roz))
    xx, xy, xz = cph * cth, cph * sth * sps - sph * cps, cph * sth * cps + sph * sps
    yx, yy, yz = sph * cth, sph * sth * sps + cph * cps, sph * sth * cps - cph * sps
    zx, zy, zz = -sth, cth * sps, cth * cps

    def mv(w, add):
        t, ns, x1, y1, z1, x2, y2, z2 = w[:8]
        p1 = (x1 * xx + y1 * xy + z1 * xz + xs, x1 * yx + y1 * yy + z1 * yz + ys, x1 * zx + y1 * zy + z1 * zz + zs)
        p2 = (x2 * xx + y2 * xy + z2 * xz + xs, x2 * yx + y2 * yy + z2 * yz + ys, x2 * zx + y2 * zy + z2 * zz + zs)
        return (t + add if (t != 0 and add) else t, ns) + p1 + p2 + tuple(w[8:])

    w0 = next((q for q, w in enumerate(wires) if its != 0 and w[0] == its), 0)
    out = list(wires)
    if nrpt == 0:
        return out[:w0] + [mv(w, 0) for w in out[w0:]]
    b0 = w0
    for _ in range(nrpt):
        b1 = len(out)
        out += [mv(out[q], itgi) for q in range(b0, b1)]
        b0 = b1
    return out


def test_gm_copies_and_moves():
    text = """CE
GW 1 3 0 0 0 0 0 0.1 0.001
GW 2 3 0 0.1 0 0 0.1 0.1 0.001
GM 5 3 0 0 30 0.2 0 0 2
GM 0 0 10 20 0 0 0 0.05 0
GE 0
EN
"""
    ctx, rc = run(text)
    w = [(1, 3, 0, 0, 0, 0, 0, 0.1, 0.001, 1.0, 1.0), (2, 3, 0, 0.1, 0, 0, 0.1, 0.1, 0.001, 1.0, 1.0)]
    w = gm(w, 5, 3, 0, 0, 30, 0.2, 0, 0, its=2)
    w = gm(w, 0, 0, 10, 20, 0, 0, 0, 0.05)
    assert len(w) == 5 and [x[0] for x in w] == [1, 2, 7, 12, 17]
    ref = api_ctx(w)
    try:
        assert rc == 0, api.nec_error_message()
        assert same_geometry(ctx, ref)
    finally:
        api.nec_delete(ctx)
        api.nec_delete(ref)
