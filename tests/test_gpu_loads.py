"""GPU parity with loads (LD cards): the HIP fill and solve through the C ABI
against the oracle and the reference's loaded-dipole known answer.  Bar as
the free-space tests: matrix entries 1e-12 relative, solutions and fields
1e-10."""
import numpy as np
import pytest
import torch  # noqa: F401  (imported at collection: the first import on a fresh box is slow)

from necpp_amd import decks
from necpp_amd import necpp as api
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

TOL = 1e-10
DIPOLE7 = [(0, 7, 0.0, 0.0, -0.25, 0.0, 0.0, 0.25, 0.001, 1.0, 1.0)]
YAGI = decks.yagi([0.5, 0.48, 0.45, 0.44], [0.4, 0.3, 0.35], 0.003, 11, rp=None)
# one card of every type; type 5 is aluminium wire everywhere (tag 0, all segments)
YAGI_LOADS = [(0, 2, 5, 7, 12.0, 2.0e-9, 8.0e-11), (1, 1, 3, 3, 300.0, 1.0e-7, 2.0e-12),
              (2, 3, 0, 0, 5.0, 1.0e-6, 0.0), (3, 4, 2, 4, 2000.0, 1.0e-5, 1.0e-13),
              (4, 2, 6, 6, 7.0, -3.0, 0.0), (5, 0, 0, 0, 3.72e7, 0.0, 0.0)]


def zin(ctx, k=0):
    return complex(api.nec_impedance_real(ctx, k), api.nec_impedance_imag(ctx, k))


def ctx_with(wires, gpflag, loads):
    ctx = api.nec_create()
    for w in wires:
        assert api.nec_wire(ctx, *w) == 0
    assert api.nec_geometry_complete(ctx, gpflag) == 0, api.nec_error_message()
    for c in loads:
        assert api.nec_ld_card(ctx, *c) == 0, api.nec_error_message()
    return ctx


def test_loaded_dipole_kat_gpu():
    # src/nec_context_tb.cpp:14-48: XQ, then LD 0 0 4 4 10 3e-9 5.3e-11, then a second solve
    ctx = ctx_with(DIPOLE7, 0, [])
    try:
        assert api.nec_ex_card(ctx, 0, 0, 4, 0, 1.0, 0.0, 0, 0, 0, 0) == 0
        assert api.nec_xq_card(ctx, 0) == 0, api.nec_error_message()
        assert api.nec_ld_card(ctx, 0, 0, 4, 4, 10.0, 3.0e-9, 5.3e-11) == 0
        assert api.nec_xq_card(ctx, 0) == 0, api.nec_error_message()
        z0, z1 = zin(ctx, 0), zin(ctx, 1)
        assert abs(z0 - complex(82.698, 46.306)) < 5e-4 * abs(z0)
        assert abs(z1.real - 92.698) < 5e-3 and abs(z1.imag - 41.941) < 5e-3, z1
        g = po.Geometry()
        g.wire(*DIPOLE7[0])
        g.complete(0)
        g.set_loads([(0, 0, 4, 4, 10.0, 3.0e-9, 5.3e-11)])
        o = g.run_voltage(po.speed_of_light() / 1e6, [(4, 1.0)])
        assert abs(z1 - o["zin"][0]) < TOL * abs(z1)
    finally:
        api.nec_delete(ctx)


def oracle_yagi(loads):
    g = po.geometry_from_deck(YAGI)
    g.set_loads(loads)
    return g


def test_fill_with_every_load_type_matches_oracle():
    ctx = ctx_with(YAGI["wires"], YAGI["gpflag"], YAGI_LOADS)
    try:
        Z = api.gpu_fill_matrix(ctx, YAGI["freq_mhz"])
        g = oracle_yagi(YAGI_LOADS)
        g.scale(YAGI["freq_mhz"])
        oZ, _ = g.fill(nthreads=4)
        err = np.abs(Z - oZ)
        bound = 1e-12 * np.abs(oZ) + 1e-15 * np.max(np.abs(oZ))
        assert np.all(err <= bound), (np.max(err / (np.abs(oZ) + 1e-300)), np.count_nonzero(err > bound))
        # the loads changed the diagonal band
        g0 = oracle_yagi([])
        g0.scale(YAGI["freq_mhz"])
        oZ0, _ = g0.fill(nthreads=4)
        assert np.max(np.abs(oZ - oZ0)) > 1e-3 * np.max(np.abs(oZ0))
    finally:
        api.nec_delete(ctx)


@pytest.mark.parametrize("precision", [0, 1])
def test_loaded_yagi_solve_and_directive_gain_gpu(precision):
    ctx = ctx_with(YAGI["wires"], YAGI["gpflag"], YAGI_LOADS)
    try:
        assert api.nec_gpu_set_precision(ctx, precision) == 0
        assert api.nec_fr_card(ctx, 0, 1, YAGI["freq_mhz"], 0.0) == 0
        s = YAGI["source"]
        assert api.nec_ex_card(ctx, 0, s[0], s[1], 0, 1.0, 0.0, 0, 0, 0, 0) == 0
        # directive gain (D = 1): the radiated power excludes the loads' loss
        assert api.nec_rp_card(ctx, 0, 19, 5, 0, 0, 1, 0, 0.0, 0.0, 10.0, 90.0, 0.0, 0.0) == 0, \
            api.nec_error_message()
        g = oracle_yagi(YAGI_LOADS)
        rp = dict(n_theta=19, n_phi=5, theta0=0.0, phi0=0.0, dtheta=10.0, dphi=90.0, ipd=1)
        o = g.run_voltage(YAGI["freq_mhz"], [(g.segment_number(*s), 1.0)], rp=rp)
        z = zin(ctx)
        assert abs(z - o["zin"][0]) < TOL * abs(z)
        cur = api.gpu_currents(ctx, 0)
        assert np.linalg.norm(cur - o["current"]) <= TOL * np.linalg.norm(o["current"])
        a = api.gpu_pattern_array(ctx, 0, "gain", 19 * 5)
        live = o["pattern"]["gain"] > -999
        assert np.max(np.abs(a[live] - o["pattern"]["gain"][live])) < 1e-8
        assert abs(api.nec_gain_max(ctx, 0) - o["stats"][0]) < 1e-9
    finally:
        api.nec_delete(ctx)


def test_ld_card_sequence_rules():
    # consecutive LD cards accumulate; after another card an LD card starts
    # a new set; LD -1 clears (src/nec_context.cpp:424-474)
    f = 299.792458
    ctx = ctx_with(DIPOLE7, 0, [])
    try:
        assert api.nec_fr_card(ctx, 0, 1, f, 0.0) == 0
        assert api.nec_ld_card(ctx, 4, 0, 4, 4, 3.0, 0.0, 0.0) == 0
        assert api.nec_ld_card(ctx, 4, 0, 4, 4, 4.0, 0.0, 0.0) == 0
        assert api.nec_ex_card(ctx, 0, 0, 4, 0, 1.0, 0.0, 0, 0, 0, 0) == 0
        assert api.nec_xq_card(ctx, 0) == 0
        z_7 = zin(ctx, 0)
        assert api.nec_ld_card(ctx, 4, 0, 4, 4, 10.0, 0.0, 0.0) == 0   # replaces both
        assert api.nec_xq_card(ctx, 0) == 0
        z_10 = zin(ctx, 1)
        assert api.nec_ex_card(ctx, 0, 0, 4, 0, 1.0, 0.0, 0, 0, 0, 0) == 0
        assert api.nec_ld_card(ctx, -1, 0, 0, 0, 0.0, 0.0, 0.0) == 0   # clears
        assert api.nec_xq_card(ctx, 0) == 0
        z_0 = zin(ctx, 2)
        assert abs((z_7 - z_0) - 7.0) < 1e-9 and abs((z_10 - z_0) - 10.0) < 1e-9, (z_0, z_7, z_10)
        # errors: no segment with the tag; the adjoint refuses loads
        assert api.nec_ld_card(ctx, 0, 9, 0, 0, 1.0, 0.0, 0.0) == 0
        assert api.nec_xq_card(ctx, 0) == 1
        assert "NO SEGMENT HAS AN ITAG" in api.nec_error_message()
        assert api.nec_ld_card(ctx, 0, 0, 4, 4, 1.0, 0.0, 0.0) == 0
        assert api.nec_xq_card(ctx, 0) == 0
        with pytest.raises(RuntimeError):
            api.gpu_adjoint_gradient(ctx)
    finally:
        api.nec_delete(ctx)
