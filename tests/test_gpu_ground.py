"""GPU parity over a ground (GN 1 perfect, GN 0 finite ground by reflection
coefficients): the HIP fill and pattern through the C ABI against the oracle
and the reference's example-3 known answers.  Bar as the free-space tests:
1e-12 relative on matrix entries, 1e-10 on solutions and fields."""
import numpy as np
import pytest
import torch  # noqa: F401  (imported at collection: the first import on a fresh box is slow)

from necpp_amd import necpp as api
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

TOL = 1e-10
GROUNDS = {"perfect": (1, 0, 0.0, 0.0), "finite": (0, 0, 6.0, 1.0e-3)}


def zin(ctx, k=0):
    return complex(api.nec_impedance_real(ctx, k), api.nec_impedance_imag(ctx, k))


def ctx_with(wires, gpflag, gn, ek=False):
    ctx = api.nec_create()
    for w in wires:
        assert api.nec_wire(ctx, *w) == 0
    assert api.nec_geometry_complete(ctx, gpflag) == 0, api.nec_error_message()
    if ek:
        assert api.nec_ek_card(ctx, 0) == 0
    if gn is not None:
        assert api.nec_gn_card(ctx, gn[0], gn[1], gn[2], gn[3], 0, 0, 0, 0) == 0, api.nec_error_message()
    return ctx


def oracle_geom(wires, gpflag, gn):
    g = po.Geometry()
    for w in wires:
        g.wire(*w)
    g.complete(gpflag)
    if gn is not None:
        g.set_ground(gn[0], gn[2], gn[3])
    return g


EX3 = [(0, 9, 0.0, 0.0, 2.0, 0.0, 0.0, 7.0, 0.03, 1.0, 1.0)]
MONOPOLE = [(1, 10, 0.0, 0.0, 0.0, 0.0, 0.0, 0.25, 0.001, 1.0, 1.0)]
# a sloping wire and a horizontal one joined above ground, one end grounded
BENT = [(1, 8, 0.0, 0.0, 0.0, 0.1, 0.0, 0.3, 0.002, 1.0, 1.0),
        (2, 9, 0.1, 0.0, 0.3, 0.45, 0.2, 0.3, 0.002, 1.0, 1.0)]


@pytest.mark.parametrize("kind", ["perfect", "finite"])
@pytest.mark.parametrize("name,wires,gpflag,f,ek", [("ex3", EX3, 1, 30.0, True),
                                                    ("bent", BENT, 1, 299.792458, False)])
def test_fill_over_ground_matches_oracle(kind, name, wires, gpflag, f, ek):
    gn = GROUNDS[kind]
    ctx = ctx_with(wires, gpflag, gn, ek)
    try:
        Z = api.gpu_fill_matrix(ctx, f)
        g = oracle_geom(wires, gpflag, gn)
        g.scale(f)
        oZ, _ = g.fill(use_ek=ek, nthreads=4)
        err = np.abs(Z - oZ)
        bound = 1e-12 * np.abs(oZ) + 1e-15 * np.max(np.abs(oZ))
        assert np.all(err <= bound), (np.max(err / (np.abs(oZ) + 1e-300)), np.count_nonzero(err > bound))
    finally:
        api.nec_delete(ctx)


@pytest.mark.parametrize("kind", ["perfect", "finite"])
def test_example3_over_ground_gpu(kats, kind):
    k = kats["example3_ground"]
    kk = k[kind]
    ctx = ctx_with(EX3, 1, GROUNDS[kind], ek=True)
    try:
        assert api.nec_fr_card(ctx, 0, 1, 30.0, 0.0) == 0
        assert api.nec_ex_card(ctx, 0, 0, 5, 0, 1.0, 0.0, 0, 0, 0, 0) == 0
        assert api.nec_rp_card(ctx, *k["deck"]["rp"]) == 0, api.nec_error_message()
        z = zin(ctx)
        tol = 0.5 * 10.0 ** (-kk["places"])
        assert abs(z.real - kk["zin"][0]) < tol and abs(z.imag - kk["zin"][1]) < tol, z
        assert abs(api.nec_gain_max(ctx, 0) - kk["gain_max"]) < tol
        g = oracle_geom(EX3, 1, GROUNDS[kind])
        rp = dict(n_theta=10, n_phi=2, theta0=0.0, phi0=0.0, dtheta=10.0, dphi=90.0, output_format=1,
                  normalization=3, ipd=0, iavp=1)
        o = g.run_voltage(30.0, [(5, 1.0)], use_ek=True, rp=rp)
        assert abs(z - o["zin"][0]) < TOL * abs(z)
        assert abs(api.nec_gain_max(ctx, 0) - o["max_gain"]) < 1e-9
        for name in ("e_theta", "e_phi"):
            a = api.gpu_pattern_array(ctx, 0, name, 20)
            assert np.max(np.abs(a - o["pattern"][name])) <= TOL * np.max(np.abs(o["pattern"][name])), name
    finally:
        api.nec_delete(ctx)


@pytest.mark.parametrize("kind", ["perfect", "finite"])
def test_grounded_monopole_full_sphere_gpu(kind):
    """Ground-connected wire end (gpflag 1), pattern over the full sphere:
    directions below the horizon read as no field on both sides."""
    gn = GROUNDS[kind]
    f = 299.792458
    ctx = ctx_with(MONOPOLE, 1, gn)
    try:
        assert api.nec_fr_card(ctx, 0, 1, f, 0.0) == 0
        assert api.nec_ex_card(ctx, 0, 1, 1, 0, 1.0, 0.0, 0, 0, 0, 0) == 0
        assert api.nec_rp_card(ctx, 0, 37, 9, 0, 0, 0, 0, 0.0, 0.0, 5.0, 45.0, 0.0, 0.0) == 0, \
            api.nec_error_message()
        g = oracle_geom(MONOPOLE, 1, gn)
        rp = dict(n_theta=37, n_phi=9, theta0=0.0, phi0=0.0, dtheta=5.0, dphi=45.0)
        o = g.run_voltage(f, [(1, 1.0)], rp=rp)
        z = zin(ctx)
        assert abs(z - o["zin"][0]) < TOL * abs(z)
        cur = api.gpu_currents(ctx, 0)
        assert np.linalg.norm(cur - o["current"]) <= TOL * np.linalg.norm(o["current"])
        npt = 37 * 9
        for name in ("gain", "pg_tot", "pg_vert", "pg_horiz"):
            a = api.gpu_pattern_array(ctx, 0, name, npt)
            b = o["pattern"][name]
            live = b > -999
            assert np.array_equal(a > -999, live), name
            # (a vertical monopole has no minor-axis field anywhere)
            if live.any():
                assert np.max(np.abs(a[live] - b[live])) < 1e-8, name
        for name in ("e_theta", "e_phi"):
            a = api.gpu_pattern_array(ctx, 0, name, npt)
            b = o["pattern"][name]
            assert np.max(np.abs(a - b)) <= TOL * np.max(np.abs(b)), name
        assert abs(api.nec_gain_max(ctx, 0) - o["max_gain"]) < 1e-9
    finally:
        api.nec_delete(ctx)


def test_mixed_precision_over_ground_gpu():
    gn = GROUNDS["finite"]
    f = 299.792458
    out = []
    for precision in (0, 1):
        ctx = ctx_with(BENT, 1, gn)
        try:
            assert api.nec_gpu_set_precision(ctx, precision) == 0
            assert api.nec_fr_card(ctx, 0, 1, f, 0.0) == 0
            assert api.nec_ex_card(ctx, 0, 2, 4, 0, 1.0, 0.0, 0, 0, 0, 0) == 0
            assert api.nec_xq_card(ctx, 0) == 0, api.nec_error_message()
            out.append(zin(ctx))
        finally:
            api.nec_delete(ctx)
    assert abs(out[1] - out[0]) < TOL * abs(out[0]), out


def test_ground_cards_outside_the_gpu_path_are_refused():
    ctx = api.nec_create()
    try:
        assert api.nec_wire(ctx, *MONOPOLE[0]) == 0
        assert api.nec_geometry_complete(ctx, 1) == 0
        assert api.nec_gn_card(ctx, 2, 0, 10.0, 0.01, 0, 0, 0, 0) == 1
        assert "Sommerfeld" in api.nec_error_message()
        assert api.nec_gn_card(ctx, 0, 8, 10.0, 0.01, 0.5, 0.001, 0, 0) == 1
        assert "radial" in api.nec_error_message()
        assert api.nec_gn_card(ctx, 0, 0, 0.0, 0.01, 0, 0, 0, 0) == 1
        # the adjoint is free-space only
        assert api.nec_gn_card(ctx, 1, 0, 0, 0, 0, 0, 0, 0) == 0
        assert api.nec_ex_card(ctx, 0, 1, 1, 0, 1.0, 0.0, 0, 0, 0, 0) == 0
        assert api.nec_xq_card(ctx, 0) == 0, api.nec_error_message()
        with pytest.raises(RuntimeError):
            api.gpu_adjoint_gradient(ctx)
    finally:
        api.nec_delete(ctx)
