"""Near electric (NE card) and magnetic (NH card) fields on the GPU against
the oracle's nefld and nhfld restatements (pinned in test_oracle_near.py):
rectangular and spherical grids, points on the wire, near it and far away,
the extended kernel, a grid structure with junctions, free space and the
perfect and reflection-coefficient grounds, and the refusals."""
import numpy as np
import pytest
import torch  # noqa: F401  (imported at collection: the first import on a fresh box is slow)

from necpp_amd import decks
from necpp_amd import necpp as api
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

C = po.speed_of_light()
TOL = 1e-10
DIPOLE = (1, 21, 0.0, 0.0, -0.25, 0.0, 0.0, 0.25, 0.001, 1.0, 1.0)


def near(ctx, k):
    cnt = api.nec_gpu_near_field(ctx, k, None, 0)
    assert cnt > 0
    out = np.zeros(9 * cnt)
    assert api.nec_gpu_near_field(ctx, k, out.ctypes.data, cnt) == cnt
    out = out.reshape(cnt, 9)
    return out[:, 0:3], out[:, 3::2] + 1j * out[:, 4::2]


def oracle_near(wires, gpflag, f, src, pts_m, use_ek=False, magnetic=False, ground=(-1, 0.0, 0.0)):
    g = po.Geometry()
    for w in wires:
        g.wire(*w)
    g.complete(gpflag)
    g.set_ground(*ground)
    o = g.run_voltage(f, [(src, 1.0)], use_ek=use_ek, nthreads=8)
    g.scale(f)
    pts = pts_m / (C / (f * 1e6))
    return g.near_h(o["abc"], pts) if magnetic else g.near_e(o["abc"], pts, use_ek=use_ek)


def check(E, Eo):
    # relative to the largest field of the grid: on the wire the field along
    # it is what is left of the cancellation the equations impose (~1e-3 of
    # the field beside the wire), so a point's own scale would amplify
    # rounding there
    err = np.max(np.abs(E - Eo)) / np.max(np.abs(Eo))
    assert err < TOL, err


@pytest.mark.parametrize("ek", [False, True])
def test_rectangular_grid_matches_oracle(ek):
    # x from the wire's surface outwards, z along the wire: the first column
    # lies on the wire (the observation radius is the wire's)
    f = 299.792458
    ctx = api.nec_create()
    try:
        assert api.nec_wire(ctx, *DIPOLE) == 0 and api.nec_geometry_complete(ctx, 0) == 0
        if ek:
            assert api.nec_ek_card(ctx, 0) == 0
        assert api.nec_fr_card(ctx, 0, 1, f, 0.0) == 0
        assert api.nec_ex_card(ctx, 0, 1, 11, 0, 1.0, 0.0, 0, 0, 0, 0) == 0
        assert api.nec_ne_card(ctx, 0, 4, 1, 6, 0.0, 0.0, -0.2, 0.02, 0.0, 0.08) == 0, api.nec_error_message()
        pts, E = near(ctx, 0)
        assert pts.shape == (24, 3) and np.allclose(pts[0], [0.0, 0.0, -0.2]) and np.allclose(pts[1], [0.02, 0, -0.2])
        check(E, oracle_near([DIPOLE], 0, f, 11, pts, use_ek=ek))
    finally:
        api.nec_delete(ctx)


def test_spherical_grid_and_a_far_point():
    f = 299.792458
    ctx = api.nec_create()
    try:
        assert api.nec_wire(ctx, *DIPOLE) == 0 and api.nec_geometry_complete(ctx, 0) == 0
        assert api.nec_fr_card(ctx, 0, 1, f, 0.0) == 0
        assert api.nec_ex_card(ctx, 0, 1, 11, 0, 1.0, 0.0, 0, 0, 0, 0) == 0
        assert api.nec_xq_card(ctx, 0) == 0
        # r = 0.6 and 200 m, theta 30..150, phi 0 and 45 (NE after XQ reuses the solve)
        assert api.nec_ne_card(ctx, 1, 2, 2, 3, 0.6, 0.0, 30.0, 199.4, 45.0, 60.0) == 0, api.nec_error_message()
        pts, E = near(ctx, 0)
        assert pts.shape == (12, 3)
        check(E, oracle_near([DIPOLE], 0, f, 11, pts))
        assert api.nec_gpu_near_field(ctx, 1, None, 0) == -1
    finally:
        api.nec_delete(ctx)


def test_grid_structure_matches_oracle():
    d = decks.grid(9, rp=None)
    f = d.get("freq_mhz", 299.792458)
    ctx = api.nec_create()
    try:
        for w in d["wires"]:
            assert api.nec_wire(ctx, *w) == 0
        assert api.nec_geometry_complete(ctx, d["gpflag"]) == 0
        assert api.nec_fr_card(ctx, 0, 1, f, 0.0) == 0
        src = d["source"]
        assert api.nec_ex_card(ctx, 0, src[0], src[1], 0, 1.0, 0.0, 0, 0, 0, 0) == 0
        assert api.nec_ne_card(ctx, 0, 5, 5, 2, -0.3, -0.3, 0.05, 0.15, 0.15, 0.2) == 0, api.nec_error_message()
        pts, E = near(ctx, 0)
        g = po.geometry_from_deck(d)
        o = g.run_voltage(f, [(g.segment_number(*src), 1.0)], nthreads=8)
        g.scale(f)
        check(E, g.near_e(o["abc"], pts / (C / (f * 1e6))))
    finally:
        api.nec_delete(ctx)


@pytest.mark.parametrize("then", ["xq", "rp"])
def test_near_field_at_every_frequency_of_a_sweep(then):
    # FR with 3 frequencies, NE, then XQ or RP: the near field is evaluated
    # at every frequency (src/nec_context.cpp:1741-1757), and after RP the
    # pattern of every frequency follows it
    freqs = (250.0, 275.0, 300.0)
    ctx = api.nec_create()
    try:
        assert api.nec_wire(ctx, *DIPOLE) == 0 and api.nec_geometry_complete(ctx, 0) == 0
        assert api.nec_fr_card(ctx, 0, 3, freqs[0], 25.0) == 0
        assert api.nec_ex_card(ctx, 0, 1, 11, 0, 1.0, 0.0, 0, 0, 0, 0) == 0
        assert api.nec_ne_card(ctx, 0, 3, 1, 4, 0.01, 0.0, -0.15, 0.03, 0.0, 0.1) == 0, api.nec_error_message()
        assert api.nec_gpu_near_field(ctx, 0, None, 0) == -1  # nothing runs before XQ / RP
        if then == "xq":
            assert api.nec_xq_card(ctx, 0) == 0, api.nec_error_message()
        else:
            assert api.nec_rp_card(ctx, 0, 5, 4, 1, 0, 0, 0, 0.0, 0.0, 45.0, 90.0, 0.0, 0.0) == 0, \
                api.nec_error_message()
        for k, f in enumerate(freqs):
            pts, E = near(ctx, k)
            assert pts.shape == (12, 3)
            check(E, oracle_near([DIPOLE], 0, f, 11, pts))
            if then == "rp":
                assert api.nec_gain_max(ctx, k) > 0.0
        assert api.nec_gpu_near_field(ctx, 3, None, 0) == -1
        # the card is spent: a later XQ adds no near field
        assert api.nec_xq_card(ctx, 0) == 0
        assert api.nec_gpu_near_field(ctx, 3, None, 0) == -1
    finally:
        api.nec_delete(ctx)


def test_magnetic_rectangular_grid_matches_oracle():
    # NH: x from the wire's surface outwards, z along the wire and beyond its
    # ends (the near-axis branch of hsflx)
    f = 299.792458
    ctx = api.nec_create()
    try:
        assert api.nec_wire(ctx, *DIPOLE) == 0 and api.nec_geometry_complete(ctx, 0) == 0
        assert api.nec_fr_card(ctx, 0, 1, f, 0.0) == 0
        assert api.nec_ex_card(ctx, 0, 1, 11, 0, 1.0, 0.0, 0, 0, 0, 0) == 0
        assert api.nec_nh_card(ctx, 0, 4, 1, 8, 0.0005, 0.0, -0.31, 0.02, 0.0, 0.09) == 0, api.nec_error_message()
        assert api.nec_gpu_near_field_kind(ctx, 0) == 1
        pts, H = near(ctx, 0)
        assert pts.shape == (32, 3)
        check(H, oracle_near([DIPOLE], 0, f, 11, pts, magnetic=True))
        # then NE on the same solve: an electric result after it
        assert api.nec_ne_card(ctx, 0, 2, 1, 2, 0.05, 0.0, 0.0, 0.05, 0.0, 0.1) == 0
        assert api.nec_gpu_near_field_kind(ctx, 1) == 0 and api.nec_gpu_near_field_kind(ctx, 2) == -1
        pts, E = near(ctx, 1)
        check(E, oracle_near([DIPOLE], 0, f, 11, pts))
    finally:
        api.nec_delete(ctx)


GROUNDS = {"perfect": (1, 0.0, 0.0), "finite": (0, 13.0, 0.005)}
ABOVE = [(1, 9, -0.24, 0.05, 0.2, 0.24, -0.05, 0.2, 0.001, 1.0, 1.0),  # tilted horizontal dipole
         (2, 7, 0.4, 0.0, 0.05, 0.4, 0.0, 0.35, 0.001, 1.0, 1.0)]      # vertical wire above the ground


@pytest.mark.parametrize("field", ["NE", "NH"])
@pytest.mark.parametrize("ground", list(GROUNDS))
def test_near_fields_over_ground_match_oracle(field, ground):
    # points above the ground, on the ground plane (z = 0) and beside both
    # wires; the image terms of efld / hsfld, with the reflection
    # coefficients of the finite ground
    f = 299.792458
    iperf, epsr, sig = GROUNDS[ground]
    ctx = api.nec_create()
    try:
        for w in ABOVE:
            assert api.nec_wire(ctx, *w) == 0
        assert api.nec_geometry_complete(ctx, 1) == 0
        assert api.nec_gn_card(ctx, iperf, 0, epsr, sig, 0, 0, 0, 0) == 0, api.nec_error_message()
        assert api.nec_fr_card(ctx, 0, 1, f, 0.0) == 0
        assert api.nec_ex_card(ctx, 0, 1, 5, 0, 1.0, 0.0, 0, 0, 0, 0) == 0
        card = api.nec_ne_card if field == "NE" else api.nec_nh_card
        assert card(ctx, 0, 5, 3, 4, -0.3, -0.2, 0.0, 0.17, 0.2, 0.1) == 0, api.nec_error_message()
        assert api.nec_gpu_near_field_kind(ctx, 0) == (field == "NH")
        pts, F = near(ctx, 0)
        assert pts.shape == (60, 3) and np.all(pts[:15, 2] == 0.0)
        check(F, oracle_near(ABOVE, 1, f, 5, pts, magnetic=field == "NH", ground=GROUNDS[ground]))
    finally:
        api.nec_delete(ctx)


def test_near_field_refusals():
    ctx = api.nec_create()
    try:
        assert api.nec_wire(ctx, 1, 11, 0, 0, 0.1, 0, 0, 0.6, 0.001, 1, 1) == 0
        assert api.nec_geometry_complete(ctx, 0) == 0
        # grounds outside the GPU path: Sommerfeld, radial-wire screens
        assert api.nec_gn_card(ctx, 2, 0, 13.0, 0.005, 0, 0, 0, 0) == 1
        assert "Sommerfeld" in api.nec_error_message()
        assert api.nec_gn_card(ctx, 0, 12, 13.0, 0.005, 10.0, 0.001, 0, 0) == 1
        assert "radial" in api.nec_error_message()
        assert api.nec_gpu_near_field_kind(ctx, 0) == -1
    finally:
        api.nec_delete(ctx)
