"""Pin the oracle's networks and transmission lines (the netwk restatement,
src/nec_context.cpp:4496-5045; TL / NT cards :630-726).  The reference's
tests hold no known answers for them, so the checks come from circuit
theory:
  * a network with both ports on one segment is a one-port of admittance
    y11 + 2 y12 + y22; across a segment's gap it carries the segment's
    current, so it is a series impedance 1 / y, i.e. the LD card's series
    load (pinned by the loaded-dipole known answers);
  * on the driven segment the one-port is in parallel with the structure:
    1 / Z_in = 1 / Z_in(without it) + y;
  * a transmission line is the two-port with the textbook Y parameters of a
    lossless line, y11 = y22 = -j cot(beta l) / Z0, y12 = j / (Z0 sin beta
    l), plus its end admittances; a crossed line negates y12; a length of 0
    means the distance of the port segments' centres;
  * the currents of a structure with two-port networks between undriven
    segments equal an independent solve of the augmented system (structure
    equations with the port voltages as gap sources, plus the port
    equations), written from those definitions with numpy."""
import numpy as np
import pytest

from oracle import pyoracle as po

C = po.speed_of_light()
F = 299.792458


def dipole_pair():
    # two parallel dipoles 0.15 wavelengths apart, 11 segments each
    g = po.Geometry()
    g.wire(1, 11, 0.0, 0.0, -0.24, 0.0, 0.0, 0.24, 0.002, 1.0, 1.0)
    g.wire(2, 11, 0.15, 0.0, -0.23, 0.15, 0.0, 0.23, 0.002, 1.0, 1.0)
    g.complete(0)
    return g


def run(g, src=((6, 1.0),), nets=(), loads=()):
    g.set_networks(list(nets))
    g.set_loads(list(loads))
    return g.run_voltage(F, list(src))


def test_one_port_across_a_gap_is_a_series_load():
    g = dipole_pair()
    z = complex(30.0, -45.0)
    y = 1.0 / z
    seg = 17  # tag 2, segment 6
    a = run(g, nets=[("NT", seg, seg, (y.real, y.imag, 0.0, 0.0, 0.0, 0.0))])
    b = run(g, loads=[(4, 2, 6, 6, z.real, z.imag, 0.0)])
    assert np.linalg.norm(a["current"] - b["current"]) < 1e-11 * np.linalg.norm(b["current"])
    assert abs(a["zin"][0] - b["zin"][0]) < 1e-11 * abs(b["zin"][0])
    # split over y11, y12 and y22 (y11 + 2 y12 + y22 = y)
    c = run(g, nets=[("NT", seg, seg, (0.5 * y.real, 0.0, 0.25 * y.real, 0.5 * y.imag, 0.0, 0.0))])
    assert np.linalg.norm(c["current"] - b["current"]) < 1e-11 * np.linalg.norm(b["current"])


def test_one_port_on_the_driven_segment_is_in_parallel():
    g = dipole_pair()
    y = complex(0.004, 0.011)
    z0 = run(g)["zin"][0]
    a = run(g, nets=[("NT", 6, 6, (y.real, y.imag, 0.0, 0.0, 0.0, 0.0))])
    assert abs(1.0 / a["zin"][0] - (1.0 / z0 + y)) < 1e-12 * abs(1.0 / z0)
    # the structure's currents are those of the source alone
    assert np.linalg.norm(a["current"] - run(g)["current"]) < 1e-13 * np.linalg.norm(a["current"])


def line_y(z0, length_m, f=F):
    bl = 2 * np.pi * length_m / (C / (f * 1e6))
    return -1j / (np.tan(bl) * z0), 1j / (np.sin(bl) * z0)


@pytest.mark.parametrize("crossed", [False, True])
def test_transmission_line_is_its_textbook_two_port(crossed):
    g = dipole_pair()
    z0, length = 75.0, 0.37
    ys1, ys2 = complex(0.001, -0.002), complex(0.0, 0.003)
    tl = ("TL", 6, 17, (-z0 if crossed else z0, length, ys1.real, ys1.imag, ys2.real, ys2.imag))
    a = run(g, nets=[tl])
    yd, yx = line_y(z0, length)
    if crossed:
        yx = -yx
    y11, y22 = yd + ys1, yd + ys2
    b = run(g, nets=[("NT", 6, 17, (y11.real, y11.imag, yx.real, yx.imag, y22.real, y22.imag))])
    assert np.linalg.norm(a["current"] - b["current"]) < 1e-10 * np.linalg.norm(b["current"])
    assert abs(a["zin"][0] - b["zin"][0]) < 1e-10 * abs(b["zin"][0])


def test_line_of_length_zero_spans_the_port_centres():
    g = dipole_pair()
    g.scale(F)
    tab = g.segments()[0]
    d = np.linalg.norm(tab[16, 0:3] - tab[5, 0:3])  # metres
    a = run(g, nets=[("TL", 6, 17, (50.0, 0.0, 0.0, 0.0, 0.0, 0.0))])
    b = run(g, nets=[("TL", 6, 17, (50.0, d, 0.0, 0.0, 0.0, 0.0))])
    assert abs(a["zin"][0] - b["zin"][0]) < 1e-10 * abs(b["zin"][0])


def augmented(g, src_seg, v, ports, Y):
    """Structure + two-port networks on undriven segments, solved at once:
    unknowns the basis amplitudes x and the port voltages V; structure rows
    Z x + sum_p V_p e_p / (s_p lambda) = -v e_src / (s_src lambda) (a port
    voltage acts as a gap source), port rows I_p + (Y V)_p = 0 with I_p =
    lambda (T x)_p the segment current (the network draws the segment's
    current: a one-port is a series impedance 1 / y)."""
    g.scale(F)
    wl = C / (F * 1e6)
    A, _ = g.fill(nthreads=4)  # row = observation segment, column = basis
    n = A.shape[0]
    T = np.zeros((n, n), complex)
    for k in range(n):
        e = np.zeros(n, complex)
        e[k] = 1.0
        T[:, k] = g.current_coefficients(e)[0]
    s = g.segments()[0][:, 3] / wl
    P = len(ports)
    M = np.zeros((n + P, n + P), complex)
    rhs = np.zeros(n + P, complex)
    M[:n, :n] = A
    for i, p in enumerate(ports):
        M[p - 1, n + i] = 1.0 / (s[p - 1] * wl)
        M[n + i, :n] = wl * T[p - 1, :]
        M[n + i, n:] = Y[i]
    rhs[src_seg - 1] = -v / (s[src_seg - 1] * wl)
    sol = np.linalg.solve(M, rhs)
    return wl * (T @ sol[:n])


def test_two_port_networks_equal_the_augmented_solve():
    g = dipole_pair()
    # a line and a network, both between undriven segments, sharing one
    y11, y12, y22 = complex(0.002, 0.004), complex(-0.001, 0.0025), complex(0.0015, -0.003)
    nets = [("NT", 3, 14, (y11.real, y11.imag, y12.real, y12.imag, y22.real, y22.imag)),
            ("TL", 14, 20, (60.0, 0.21, 0.0, 0.0, 0.0, 0.0))]
    a = run(g, nets=nets)
    yd, yx = line_y(60.0, 0.21)
    ports = [3, 14, 20]
    Y = np.zeros((3, 3), complex)
    Y[0, 0], Y[0, 1], Y[1, 0], Y[1, 1] = y11, y12, y12, y22
    Y[1, 1] += yd
    Y[1, 2] += yx
    Y[2, 1] += yx
    Y[2, 2] += yd
    want = augmented(g, 6, 1.0, ports, Y)
    assert np.linalg.norm(a["current"] - want) < 1e-10 * np.linalg.norm(want)


def test_source_missing_from_the_network_table_is_an_error():
    # a network on one of two sources: the other source is not in the
    # network's table and the reference stops (NETWORK DATA ERROR)
    g = dipole_pair()
    with pytest.raises(RuntimeError, match="NETWORK DATA ERROR"):
        run(g, src=((6, 1.0), (17, 1.0)), nets=[("NT", 6, 6, (0.01, 0.0, 0.0, 0.0, 0.0, 0.0))])
