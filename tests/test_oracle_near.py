"""Pin the oracle's near fields (the nefld and nhfld restatements) by what
the mathematics fixes, since the reference's own tests check only the number
of near-field points: (1) at every segment centre the electric field along
the segment cancels the incident field the equations were solved for (the
self segment is integrated as a neighbour there, so the agreement is to the
adaptive quadrature's accuracy); (2) far from the structure it is the far
field of the pattern at that range; (3) the magnetic field is the curl of
the electric one over -j omega mu0, circles a wire as I / (2 pi rho) next to
it and is r x E / eta far away; (4) over a perfect ground the tangential E
and the normal H vanish on the ground plane, and the reflection-coefficient
ground tends to the perfect ground and to free space in its limits."""
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


# ------------------------------------------------------------ magnetic field
# The oracle's nhfld restatement (hsfld, hsflx, hfk) is pinned by Maxwell's
# equations against the electric field pinned above, by the magnetostatic
# limit next to a wire, and by the plane-wave relation far away; its ground
# terms by the boundary conditions on a perfect ground and by the limits of
# the finite ground.

ETA = po.impedance()
MU0 = ETA / C  # eta = mu0 c


def curl_e(g, abc, p, wl, h=2e-3):
    """curl E at p (wavelengths) by fourth-order central differences of the
    oracle's E, in V/m per metre."""
    def E(q):
        return g.near_e(abc, np.asarray(q, float).reshape(1, 3))[0]
    d = np.zeros((3, 3), complex)  # d[a][b] = dE_b / dx_a
    for a in range(3):
        e = np.zeros(3)
        e[a] = h
        d[a] = (8 * (E(p + e) - E(p - e)) - (E(p + 2 * e) - E(p - 2 * e))) / (12 * h * wl)
    return np.array([d[1][2] - d[2][1], d[2][0] - d[0][2], d[0][1] - d[1][0]])


def solved(g, f, seg, rp=None):
    o = g.run_voltage(f, [(seg, 1.0)], rp=rp)
    g.scale(f)
    return o


def test_faraday_curl_e_is_minus_j_omega_mu_h():
    # curl E = -j omega mu0 H (time factor exp(+j omega t), as the outgoing
    # exp(-j k r) of the kernels), at points from 0.05 to 2 wavelengths, on
    # and off the dipole's axis and beyond its ends.  Both fields carry the
    # adaptive quadratures' error (tolerance 1e-4 per step: ~1e-5 of the
    # field), and beyond rkh = 1 wavelength from a segment E treats it as a
    # lumped element (error ~ (length / r)^2 ~ 2e-4 here) where H does not
    f = 299.792458
    wl = C / (f * 1e6)
    w = 2 * np.pi * f * 1e6
    g = dipole()
    o = solved(g, f, 11)
    pts = np.array([[0.05, 0.0, 0.0], [0.1, 0.07, 0.1], [0.3, -0.2, 0.4], [0.0, 0.2, -0.45], [1.2, 0.9, -0.7]])
    bars = [1e-4, 1e-4, 1e-4, 1e-4, 1e-3]
    H = g.near_h(o["abc"], pts)
    for p, h, bar in zip(pts, H, bars):
        want = curl_e(g, o["abc"], p, wl) / (-1j * w * MU0)
        assert np.linalg.norm(h - want) < bar * np.linalg.norm(want), (p, h, want)


def test_magnetostatic_limit_next_to_the_wire():
    # at rho = 2 radii from the middle of a segment, rho << segment length
    # << wavelength, H circles the wire as I / (2 pi rho) in the sense of the
    # current (right-hand rule)
    f = 299.792458
    wl = C / (f * 1e6)
    g = dipole(ns=21, half=0.25, rad=0.0005)
    o = solved(g, f, 11)
    i6 = o["current"][5]  # segment 6, centre z = -0.25 + 5.5 / 21 * 0.5
    z6 = (-0.25 + 5.5 * 0.5 / 21) / wl
    rho = 0.001 / wl
    for phi in (0.0, 1.0, 2.5):
        p = np.array([[rho * np.cos(phi), rho * np.sin(phi), z6]])
        h = g.near_h(o["abc"], p)[0]
        phat = np.array([-np.sin(phi), np.cos(phi), 0.0])
        want = i6 / (2 * np.pi * rho * wl)
        assert abs(np.dot(h, phat) - want) < 3e-3 * abs(want), (np.dot(h, phat), want)
        assert abs(h[2]) < 1e-6 * abs(want)


def test_far_range_h_is_r_cross_e_over_eta():
    f = 299.792458
    wl = C / (f * 1e6)
    g = dipole()
    o = solved(g, f, 11)
    for rhat in (np.array([1.0, 0.0, 0.0]), np.array([0.6, 0.0, 0.8]), np.array([0.36, 0.48, 0.8])):
        p = (300.0 * rhat).reshape(1, 3)
        E = g.near_e(o["abc"], p)[0]
        H = g.near_h(o["abc"], p)[0]
        want = np.cross(rhat, E) / ETA
        assert np.linalg.norm(H - want) < 5e-3 * np.linalg.norm(want), (H, want)


def horizontal_dipole(height=0.2, ns=15):
    g = po.Geometry()
    g.wire(1, ns, -0.24, 0.0, height, 0.24, 0.0, height, 0.001, 1.0, 1.0)
    g.complete(0)
    return g


def test_perfect_ground_boundary_conditions():
    # on a perfect ground plane (z = 0) the tangential E and the normal H
    # vanish: the image cancels the direct field there
    f = 299.792458
    g = horizontal_dipole()
    g.set_ground(1)
    o = solved(g, f, 8)
    pts = np.array([[0.0, 0.0, 0.0], [0.1, 0.05, 0.0], [-0.3, 0.2, 0.0], [0.7, -0.4, 0.0]])
    E = g.near_e(o["abc"], pts)
    H = g.near_h(o["abc"], pts)
    scale_e = np.max(np.abs(E))
    scale_h = np.max(np.abs(H))
    assert np.max(np.abs(E[:, 0:2])) < 1e-10 * scale_e
    assert np.max(np.abs(H[:, 2])) < 1e-10 * scale_h
    # and the fields there are not zero: E normal (but below the centre,
    # where the two halves' normal fields cancel by symmetry), H tangential
    assert np.min(np.abs(E[1:, 2])) > 1e-3 * scale_e and np.min(np.abs(H[:, 0]) + np.abs(H[:, 1])) > 1e-3 * scale_h
    # above the ground the image adds a field of the same order
    p = np.array([[0.1, 0.3, 0.5]])
    g.set_ground(-1)
    g.scale(f)
    E0, H0 = g.near_e(o["abc"], p)[0], g.near_h(o["abc"], p)[0]
    g.set_ground(1)
    g.scale(f)
    E1, H1 = g.near_e(o["abc"], p)[0], g.near_h(o["abc"], p)[0]
    assert np.linalg.norm(E1 - E0) > 0.1 * np.linalg.norm(E0)
    assert np.linalg.norm(H1 - H0) > 0.1 * np.linalg.norm(H0)


def test_finite_ground_limits():
    # the reflection-coefficient ground tends to the perfect ground as the
    # conductivity grows and to free space as epsr -> 1, sigma -> 0 (the
    # same current coefficients evaluated under each ground)
    f = 299.792458
    g = horizontal_dipole()
    o = solved(g, f, 8)
    pts = np.array([[0.1, 0.3, 0.5], [0.4, -0.2, 0.05], [0.0, 0.0, 0.1], [1.5, 0.5, 0.7]])

    def fields(iperf, epsr=0.0, sig=0.0):
        g.set_ground(iperf, epsr, sig)
        g.scale(f)
        return g.near_e(o["abc"], pts), g.near_h(o["abc"], pts)

    Ep, Hp = fields(1)
    Ef, Hf = fields(0, 10.0, 1e12)
    assert np.max(np.abs(Ef - Ep)) < 1e-5 * np.max(np.abs(Ep))
    assert np.max(np.abs(Hf - Hp)) < 1e-5 * np.max(np.abs(Hp))
    E0, H0 = fields(-1)
    Ev, Hv = fields(0, 1.0, 1e-20)
    assert np.max(np.abs(Ev - E0)) < 1e-9 * np.max(np.abs(E0))
    assert np.max(np.abs(Hv - H0)) < 1e-9 * np.max(np.abs(H0))
    # an ordinary ground lies between: distinct from both
    Eg, Hg = fields(0, 13.0, 0.005)
    assert np.max(np.abs(Eg - Ep)) > 1e-2 * np.max(np.abs(Ep)) and np.max(np.abs(Eg - E0)) > 1e-2 * np.max(np.abs(E0))


def test_far_range_h_over_finite_ground_pairs_with_e():
    # far above a finite ground the direct and the reflected waves travel
    # almost parallel, so H = r x E / eta holds for their sum too (to the
    # angle between them, 2 height / range): this pairs nhfld's reflection
    # coefficients (rrv, rrh on the field's components) with efld's (refs,
    # refps), which the finite-ground known answers pin
    f = 299.792458
    g = horizontal_dipole()
    g.set_ground(0, 13.0, 0.005)
    o = solved(g, f, 8)
    for rhat in (np.array([0.6, 0.0, 0.8]), np.array([0.5, 0.5, 0.70710678]), np.array([0.0, 0.9, 0.43588989])):
        p = (400.0 * rhat).reshape(1, 3)
        E = g.near_e(o["abc"], p)[0]
        H = g.near_h(o["abc"], p)[0]
        want = np.cross(rhat, E) / ETA
        assert np.linalg.norm(H - want) < 5e-3 * np.linalg.norm(want), (rhat, H, want)
