"""Pin the oracle's polarisation outputs (ellipse, sense, RHCP/LHCP split and
their statistics), which no reference test or known answer covers.

The oracle restates nec_radiation_pattern::analyze
(src/nec_radiation_pattern.cpp:289-327 ellipse and sense, :417-426 the
RHCP/LHCP split) and the statistics of src/nec_radiation_pattern.h:252-333.
These tests check them against what does not depend on that code:

* Stokes parameters of the oracle's own E_theta, E_phi at every direction:
  major/minor axes, tilt and the handedness come from I, Q, U, V, and the
  circular powers from the projections E_R = (E_theta + j E_phi)/sqrt(2),
  E_L = (E_theta - j E_phi)/sqrt(2).  For the exp(+jwt) time convention a
  right-hand (IEEE) wave travelling along r has E proportional to
  theta_hat - j phi_hat, so E_phi lagging E_theta is right-handed.
* Physics: a z-directed dipole radiates linear polarisation (equal RHCP and
  LHCP, each half the total power); a turnstile (crossed dipoles fed in
  quadrature) radiates nearly circular polarisation along its axis, with
  the sense set by the feed phase, and reversing the feed phase exchanges
  the RHCP and LHCP patterns.
"""
import numpy as np
import pytest

from necpp_amd import decks
from oracle import pyoracle as po

DB3 = 10.0 * np.log10(2.0)


RP_GRID = dict(n_theta=19, n_phi=37, theta0=0.0, phi0=0.0, dtheta=10.0, dphi=10.0, iavp=0)


def run_multi(deck, rp, nthreads=4):
    g = po.geometry_from_deck(deck)
    src = [(g.segment_number(*ts), v) for ts, v in deck["sources"]]
    res = g.run_voltage(deck["freq_mhz"], src, rp=rp, nthreads=nthreads)
    res["geometry"] = g
    return res


def angles(rp):
    th = rp["theta0"] + rp["dtheta"] * np.arange(rp["n_theta"])
    ph = rp["phi0"] + rp["dphi"] * np.arange(rp["n_phi"])
    # points are ordered phi_index * n_theta + theta_index
    return np.tile(th, rp["n_phi"]), np.repeat(ph, rp["n_theta"])


def stokes(et, ep):
    I = np.abs(et) ** 2 + np.abs(ep) ** 2
    Q = np.abs(et) ** 2 - np.abs(ep) ** 2
    U = 2.0 * np.real(et * np.conj(ep))
    V = 2.0 * np.imag(np.conj(et) * ep)  # < 0: E_phi lags E_theta, right-handed
    L = np.hypot(Q, U)
    major2 = 0.5 * (I + L)
    minor2 = np.maximum(0.5 * (I - L), 0.0)
    return I, Q, U, V, L, major2, minor2


def check_against_stokes(pat, rp):
    et, ep = pat["e_theta"], pat["e_phi"]
    I, Q, U, V, L, major2, minor2 = stokes(et, ep)
    # directions above the dB floor (db10 gives -999.99 below 1e-20)
    ok = (I > 1e-30) & (pat["pg_tot"] > -999.0)
    tot_lin = 10.0 ** (pat["pg_tot"] / 10.0)
    # axial ratio minor/major (the ellipse's minor axis squared is a
    # difference of nearly equal terms for nearly linear fields: 1e-7 absolute)
    ar = np.sqrt(minor2[ok] / major2[ok])
    assert np.max(np.abs(pat["axial_ratio"][ok] - ar)) < 1e-7
    # circular powers as fractions of the total gain (linear scale)
    er2 = np.abs(et + 1j * ep) ** 2 / 2.0
    el2 = np.abs(et - 1j * ep) ** 2 / 2.0
    rh_lin = 10.0 ** (pat["pg_rhcp"] / 10.0)
    lh_lin = 10.0 ** (pat["pg_lhcp"] / 10.0)
    assert np.max(np.abs(rh_lin[ok] - tot_lin[ok] * er2[ok] / I[ok]) / tot_lin[ok]) < 1e-9
    assert np.max(np.abs(lh_lin[ok] - tot_lin[ok] * el2[ok] / I[ok]) / tot_lin[ok]) < 1e-9
    # handedness from Stokes V where the ellipse is not degenerate
    s = pat["sense"]
    lin = ok & (pat["axial_ratio"] <= 1e-5)
    assert np.all(s[lin] == 0)
    ell = ok & (pat["axial_ratio"] > 2e-5)
    assert np.all(s[ell & (V < 0)] == 1)
    assert np.all(s[ell & (V > 0)] == 2)
    # tilt of the major axis from theta_hat toward phi_hat, where it is
    # defined (not circular), modulo 180 degrees
    tilt = np.degrees(0.5 * np.arctan2(U, Q))
    defined = ok & (L > 1e-3 * I)
    d = (pat["tilt"][defined] - tilt[defined] + 90.0) % 180.0 - 90.0
    assert np.max(np.abs(d)) < 1e-7
    # major / minor axis gains (output format 0) relative to the total
    if rp.get("output_format", 0) != 1:
        mj = 10.0 ** (pat["pg_vert"] / 10.0)
        mn = 10.0 ** (pat["pg_horiz"] / 10.0)
        assert np.max(np.abs(mj[ok] - tot_lin[ok] * major2[ok] / I[ok]) / tot_lin[ok]) < 1e-9
        assert np.max(np.abs(mn[ok] - tot_lin[ok] * minor2[ok] / I[ok]) / tot_lin[ok]) < 1e-9
    else:
        v = 10.0 ** (pat["pg_vert"] / 10.0)
        h = 10.0 ** (pat["pg_horiz"] / 10.0)
        assert np.max(np.abs(v[ok] - tot_lin[ok] * np.abs(et[ok]) ** 2 / I[ok]) / tot_lin[ok]) < 1e-9
        assert np.max(np.abs(h[ok] - tot_lin[ok] * np.abs(ep[ok]) ** 2 / I[ok]) / tot_lin[ok]) < 1e-9


def weighted_stats(p, rp):
    """max, min, mean, sd of a pattern array with the sin(theta) weights of
    the reference's statistics (rows theta, columns phi)."""
    nth, nph = rp["n_theta"], rp["n_phi"]
    th, _ = angles(rp)
    # the reference converts degrees with its pi = 3.1415926536 (math_util.h)
    w = np.sin(th * (3.1415926536 / 180.0)) * (nth - 1) / nth + 1.0 / nth
    norm = (nth * nph * 2) / 3.1415926536
    mean = np.sum(p * w) / norm
    sd = np.sqrt(np.sum((p - mean) ** 2 * w) / norm)
    return [np.max(p), np.min(p), mean, sd]


def test_linear_dipole_splits_power_equally():
    deck = decks.dipole(21, rp=None)
    rp = dict(RP_GRID, output_format=1)
    o = po.geometry_from_deck(deck)
    res = o.run_voltage(deck["freq_mhz"], [(o.segment_number(1, 11), 1.0)], rp=rp, nthreads=4)
    pat = res["pattern"]
    th, _ = angles(rp)
    on = (th > 0.5) & (th < 179.5)  # off the axis, where the field vanishes
    assert np.all(pat["axial_ratio"] == 0.0)
    assert np.all(pat["sense"][on] == 0)
    assert np.all(pat["tilt"][on] == 0.0)
    for k in ("pg_rhcp", "pg_lhcp"):
        assert np.max(np.abs(pat[k][on] - (pat["pg_tot"][on] - DB3))) < 1e-12
    check_against_stokes(pat, rp)


@pytest.mark.parametrize("fmt", [0, 1])
def test_turnstile_ellipse_against_stokes(fmt):
    rp = dict(RP_GRID, output_format=fmt)
    res = run_multi(decks.turnstile(-1j), rp)
    pat = res["pattern"]
    check_against_stokes(pat, rp)
    # the pattern is elliptical almost everywhere: a real test of the ellipse
    assert np.mean(pat["axial_ratio"] > 0.05) > 0.5


def test_turnstile_sense_follows_feed_phase():
    rp = dict(RP_GRID, output_format=0)
    r = run_multi(decks.turnstile(-1j), rp)["pattern"]
    l = run_multi(decks.turnstile(+1j), rp)["pattern"]
    th, ph = angles(rp)
    up = (th == 0.0)
    down = (th == 180.0)
    # E_phi lags E_theta on +z for the -j feed: right-handed, nearly circular
    assert np.all(r["sense"][up] == 1) and np.all(l["sense"][up] == 2)
    assert np.all(r["axial_ratio"][up] > 0.6) and np.all(l["axial_ratio"][up] > 0.6)
    assert np.all(r["pg_rhcp"][up] > r["pg_lhcp"][up] + 10.0)
    assert np.all(l["pg_lhcp"][up] > l["pg_rhcp"][up] + 10.0)
    # the wave leaving along -z turns the same way in space, so its
    # handedness relative to its own direction is the opposite
    assert np.all(r["sense"][down] == 2) and np.all(l["sense"][down] == 1)
    # the mirror y -> -y maps the structure onto itself with the y-dipole's
    # current reversed, i.e. the -j feed onto the +j feed, and a direction
    # (theta, phi) onto (theta, -phi); a mirror image turns the other way,
    # so RHCP and LHCP exchange there and the total and axial ratio stay
    nth, nph = rp["n_theta"], rp["n_phi"]
    mirror = np.array([((nph - 1 - (i // nth)) % nph) * nth + i % nth for i in range(nth * nph)])
    assert np.all(ph[mirror] == (360.0 - ph) % 360.0 + 360.0 * (ph == 0.0))
    # (the reference's pi = 3.1415926536 puts 350 and -10 degrees 1e-8 rad
    # apart, so powers are compared relative to the total, not in dB)
    finite = r["pg_tot"] > -999.0
    lin = lambda a: 10.0 ** (a / 10.0)
    tot = lin(r["pg_tot"])[finite]
    for a, b in (("pg_rhcp", "pg_lhcp"), ("pg_lhcp", "pg_rhcp"), ("pg_tot", "pg_tot")):
        assert np.max(np.abs(lin(r[a]) - lin(l[b])[mirror])[finite] / tot) < 1e-8, a
    assert np.max(np.abs(r["axial_ratio"] - l["axial_ratio"][mirror])[finite]) < 1e-7
    circ = finite & (r["axial_ratio"] > 1e-3)
    assert np.all(r["sense"][circ] + l["sense"][mirror][circ] == 3)


def test_statistics_are_over_their_own_arrays():
    # upper hemisphere only: there the -j feed is mostly right-handed, so the
    # RHCP and LHCP statistics differ and a swapped array shows
    rp = dict(RP_GRID, output_format=0, dtheta=5.0)
    res = run_multi(decks.turnstile(-1j), rp)
    pat, st = res["pattern"], res["stats"]
    # stats: max, min, mean, sd of total, RHCP, LHCP gain in that order
    for a, name in enumerate(("pg_tot", "pg_rhcp", "pg_lhcp")):
        want = weighted_stats(pat[name], rp)
        assert np.allclose(st[4 * a:4 * a + 4], want, rtol=1e-12, atol=1e-12), name
    assert st[4] > st[8] + 3.0 and st[6] > st[10] + 3.0
