"""Pin the oracle's loads (LD cards): the reference's loaded-dipole known
answer, the wire-conductivity impedance against the exact Bessel-function
formula, and identities between the load types."""
import numpy as np
import pytest
import scipy.special as sps

from oracle import pyoracle as po

TPCMU = 2.368705e3  # 2 pi c mu0 as the reference writes it (nec_context.cpp:6346)


def dipole7():
    g = po.Geometry()
    g.wire(0, 7, 0.0, 0.0, -0.25, 0.0, 0.0, 0.25, 0.001, 1.0, 1.0)
    g.complete(0)
    return g


def test_loaded_dipole_kat():
    # src/nec_context_tb.cpp:14-48: Z_in before and after LD 0 0 4 4 10 3e-9 5.3e-11
    g = dipole7()
    f = po.speed_of_light() / 1e6
    z0 = g.run_voltage(f, [(4, 1.0)])["zin"][0]
    assert abs(z0 - complex(82.698, 46.306)) < 5e-4 * abs(z0)
    g.set_loads([(0, 0, 4, 4, 10.0, 3.0e-9, 5.3e-11)])
    z1 = g.run_voltage(f, [(4, 1.0)])["zin"][0]
    assert abs(z1.real - 92.698) < 5e-3 and abs(z1.imag - 41.941) < 5e-3, z1


def exact_zint(sigl, rolam):
    # internal impedance of a round wire, (gamma / (2 pi a sigma)) J0/J1, in
    # the reference's scaled variables (times wavelength)
    xa = np.sqrt(TPCMU * sigl) * rolam
    ga = xa * (1 - 1j) / np.sqrt(2)
    return (1 - 1j) / np.sqrt(2) * TPCMU / (2 * 3.1415926536 * xa) * sps.jv(0, ga) / sps.jv(1, ga)


@pytest.mark.parametrize("x", [0.01, 0.3, 1.0, 3.0, 7.9, 8.1, 20.0, 60.0, 105.0])
def test_zint_matches_bessel_formula(x):
    rolam = 1e-4
    sigl = (x / rolam) ** 2 / TPCMU
    z = po.zint(sigl, rolam)
    ze = exact_zint(sigl, rolam)
    assert abs(z - ze) <= 2e-7 * abs(ze), (z, ze)


def test_zint_skin_effect_limit():
    # x > 110: R = X = sqrt(60 / sigl) / (sqrt 2 rolam)
    rolam, sigl = 1e-2, 1e6
    assert np.sqrt(TPCMU * sigl) * rolam > 110
    z = po.zint(sigl, rolam)
    r = np.sqrt(60.0 / sigl) * 0.70710678 / rolam
    assert abs(z.real - r) < 1e-12 * r and abs(z.imag - r) < 1e-12 * r


def run_loaded(cards, f=299.792458, rp=None):
    g = dipole7()
    g.set_loads(cards)
    return g.run_voltage(f, [(4, 1.0)], rp=rp)


def test_load_type_identities():
    f = 299.792458
    seg_len = 0.5 / 7  # metres
    base = run_loaded([(0, 0, 4, 4, 10.0, 0.0, 0.0)], f)["zin"][0]
    # a fixed impedance of 10 ohm, the same resistance per metre, a parallel
    # circuit of the resistance alone
    for cards in ([(4, 0, 4, 4, 10.0, 0.0, 0.0)], [(2, 0, 4, 4, 10.0 / seg_len, 0.0, 0.0)],
                  [(1, 0, 4, 4, 10.0, 0.0, 0.0)]):
        z = run_loaded(cards, f)["zin"][0]
        assert abs(z - base) < 1e-12 * abs(base), (cards, z, base)
    # the resistance adds to Z_in when it sits on the fed segment
    z_free = dipole7().run_voltage(f, [(4, 1.0)])["zin"][0]
    assert abs(base - (z_free + 10.0)) < 1e-9 * abs(base)
    # two cards on one segment add
    z2 = run_loaded([(4, 0, 4, 4, 4.0, 1.0, 0.0), (4, 0, 4, 4, 6.0, -1.0, 0.0)], f)["zin"][0]
    assert abs(z2 - base) < 1e-12 * abs(base)


def test_loss_lowers_power_gain_not_directive_gain():
    # the pattern's gcon uses the radiated power (input power less the
    # loads' loss) when ipd = 1: directive gain - power gain = 10 log10(pin/prad)
    rp0 = dict(n_theta=7, n_phi=1, theta0=0.0, phi0=0.0, dtheta=15.0, dphi=0.0, ipd=0)
    rp1 = dict(rp0, ipd=1)
    r0 = run_loaded([(0, 0, 4, 4, 10.0, 0.0, 0.0)], rp=rp0)
    r1 = run_loaded([(0, 0, 4, 4, 10.0, 0.0, 0.0)], rp=rp1)
    z = r0["zin"][0]
    i_src = 1.0 / z
    pin = 0.5 * (1.0 / z).real
    ploss = 0.5 * abs(i_src) ** 2 * 10.0
    assert abs((r1["max_gain"] - r0["max_gain"]) - 10 * np.log10(pin / (pin - ploss))) < 1e-9


def test_load_errors():
    g = dipole7()
    with pytest.raises(RuntimeError):
        g.set_loads([(0, 0, 5, 4, 10.0, 0.0, 0.0)])  # step1 > step2
    g.set_loads([(0, 9, 0, 0, 10.0, 0.0, 0.0)])     # no segment has tag 9
    with pytest.raises(RuntimeError):
        g.run_voltage(299.792458, [(4, 1.0)])
