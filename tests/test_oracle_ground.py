"""Pin the oracle's ground branches (GN 1 perfect ground, GN 0 finite ground
by reflection coefficients): the reference's example-3 known answers, and
image theory for what those leave open (a monopole on a perfect ground is one
half of a dipole fed at its centre by two equal sources; its pattern over
the upper half-space is the dipole's, 3.01 dB up because the source power
halves).
"""
import numpy as np
import pytest

from oracle import pyoracle as po


def ex3(kats, kind):
    k = kats["example3_ground"]
    g = po.Geometry()
    g.wire(*k["deck"]["wire"])
    g.complete(k["deck"]["gpflag"])
    gn = k[kind]["gn"]
    g.set_ground(gn[0], gn[2], gn[3])
    rp = dict(n_theta=10, n_phi=2, theta0=0.0, phi0=0.0, dtheta=10.0, dphi=90.0, output_format=1,
              normalization=3, ipd=0, iavp=1)
    return k[kind], g.run_voltage(30.0, [(5, 1.0)], use_ek=True, rp=rp)


@pytest.mark.parametrize("kind", ["perfect", "finite"])
def test_example3_ground_kats(kats, kind):
    k, r = ex3(kats, kind)
    tol = 0.5 * 10.0 ** (-k["places"])
    z = r["zin"][0]
    assert abs(z.real - k["zin"][0]) < tol and abs(z.imag - k["zin"][1]) < tol, z
    assert abs(r["max_gain"] - k["gain_max"]) < tol, r["max_gain"]
    if kind == "perfect":
        # 7 places of the test are far looser than the restatement: same
        # operations as the reference up to the LU's rounding
        assert abs(z - complex(*k["zin"])) < 1e-12 * abs(z)
        assert abs(r["max_gain"] - k["gain_max"]) < 1e-12


def monopole_and_dipole(ek):
    f = 299.792458
    rp = dict(n_theta=19, n_phi=1, theta0=0.0, phi0=0.0, dtheta=5.0, dphi=0.0)
    m = po.Geometry()
    m.wire(1, 10, 0, 0, 0.0, 0, 0, 0.25, 0.001, 1, 1)
    m.complete(1)
    m.set_ground(1)
    rm = m.run_voltage(f, [(1, 1.0)], use_ek=ek, rp=rp)
    d = po.Geometry()
    d.wire(1, 20, 0, 0, -0.25, 0, 0, 0.25, 0.001, 1, 1)
    d.complete(0)
    rd = d.run_voltage(f, [(10, 1.0), (11, 1.0)], use_ek=ek, rp=rp)
    return rm, rd


@pytest.mark.parametrize("ek", [False, True])
def test_monopole_on_perfect_ground_is_half_a_dipole(ek):
    rm, rd = monopole_and_dipole(ek)
    zm = rm["zin"][0]
    # the two sources of the symmetric dipole see the same impedance
    assert abs(rd["zin"][0] - rd["zin"][1]) < 1e-8 * abs(rd["zin"][0])
    # image theory; the residue is the adaptive quadrature running the image
    # segment in the opposite direction
    assert abs(zm - rd["zin"][0]) < 1e-7 * abs(zm), (zm, rd["zin"])
    # currents: monopole segment k is dipole segment 10 + k
    assert np.allclose(rm["current"], rd["current"][10:], rtol=1e-7, atol=1e-7 * np.abs(rd["current"]).max())
    # pattern over the upper half-space (theta 0..90): same fields, half the
    # input power -> +10 log10(2) dB
    gm = rm["pattern"]["pg_tot"]
    gd = rd["pattern"]["pg_tot"]
    live = gd > -999
    assert np.all(np.abs(gm[live] - (gd[live] + 10.0 * np.log10(2.0))) < 1e-6)


def test_pattern_below_horizon_over_ground_reads_as_no_field():
    g = po.Geometry()
    g.wire(1, 9, 0, 0, 0.2, 0, 0, 0.7, 0.003, 1, 1)
    g.complete(1)
    g.set_ground(1)
    rp = dict(n_theta=19, n_phi=2, theta0=0.0, phi0=0.0, dtheta=10.0, dphi=90.0)
    r = g.run_voltage(299.792458, [(5, 1.0)], rp=rp)
    th = np.tile(np.arange(19) * 10.0, 2)
    below = th > 90.01
    p = r["pattern"]
    # (theta = 0 is the vertical wire's own null)
    assert np.all(p["gain"][below] == -999.99) and np.all(p["gain"][~below & (th > 0)] > -999)
    assert np.all(p["e_theta"][below] == 0) and np.all(p["sense"][below] == 3)
    assert r["max_gain"] == p["gain"][~below].max()


def test_finite_ground_approaches_perfect_ground_for_a_good_conductor():
    # |zrati| -> 0: the reflection coefficients tend to the perfect ground's
    g = po.Geometry()
    g.wire(0, 9, 0.0, 0.0, 2.0, 0.0, 0.0, 7.0, 0.03, 1.0, 1.0)
    g.complete(1)
    g.set_ground(1)
    zp = g.run_voltage(30.0, [(5, 1.0)])["zin"][0]
    g.set_ground(0, 1.0, 1.0e7)
    zf = g.run_voltage(30.0, [(5, 1.0)])["zin"][0]
    assert abs(zf - zp) < 1e-4 * abs(zp), (zf, zp)


def test_ground_types_outside_the_restatement_are_refused():
    g = po.Geometry()
    with pytest.raises(RuntimeError):
        g.set_ground(2, 10.0, 0.01)
    with pytest.raises(RuntimeError):
        g.set_ground(0, 0.0, 0.01)
