"""Pin the oracle's near electric field (nefld restatement) by what the
mathematics fixes, since the reference's own tests check only the number of
near-field points: (1) at every segment centre the field along the segment
cancels the incident field the equations were solved for (the self segment
is integrated as a neighbour there, so the agreement is to the adaptive
quadrature's accuracy); (2) far from the structure it is the far field of
the pattern at that range."""
import numpy as np

from oracle import pyoracle as po

C = po.speed_of_light()


def dipole(ns=21, half=0.25, rad=0.001):
    g = po.Geometry()
    g.wire(1, ns, 0.0, 0.0, -half, 0.0, 0.0, half, rad, 1.0, 1.0)
    g.complete(0)
    return g


def test_match_points_cancel_the_incident_field():
    # the thin-wire equations make the scattered field's component along each
    # segment, at its centre, equal to minus the incident one: -V / length on
    # the driven segment (the excitation of the matrix equation) and zero on
    # the others
    f = 299.792458
    g = dipole()
    o = g.run_voltage(f, [(11, 1.0)])
    g.scale(f)
    tab, _, _, _ = g.segments()
    wl = C / (f * 1e6)
    pts = tab[:, 0:3] / wl  # segment centres (the table holds metres)
    E = g.near_e(o["abc"], pts)
    proj = np.sum(E * tab[:, 5:8], axis=1)
    e = np.zeros(len(proj), complex)
    e[10] = -1.0 / tab[10, 3]
    err = np.abs(proj - e) / abs(e[10])
    assert np.max(err) < 1e-4, np.max(err)


def test_far_range_matches_the_pattern_field():
    f = 299.792458
    wl = C / (f * 1e6)
    g = dipole()
    R = 200.0 * wl
    rp = dict(n_theta=1, n_phi=1, theta0=90.0, phi0=0.0, dtheta=0.0, dphi=0.0, output_format=1, rfld=R)
    o = g.run_voltage(f, [(11, 1.0)], rp=rp)
    g.scale(f)
    E = g.near_e(o["abc"], np.array([[R / wl, 0.0, 0.0]]))[0]
    # at theta = 90 deg, phi = 0: E_theta = -E_z
    et = o["pattern"]["e_theta"][0]
    assert abs(-E[2] - et) < 5e-3 * abs(et), (-E[2], et)
    assert abs(E[0]) < 1e-2 * abs(et) and abs(E[1]) < 1e-6 * abs(et)
