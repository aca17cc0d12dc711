"""Pin the oracle's GM (move / rotate / copy, c_geometry::move,
src/c_geometry.cpp:830-900) and GS (c_geometry::scale, :1308-1318)
restatements by closed forms: rotations by right angles and known angles
of straight wires, translations, the copy count and tag sequence, the ITS
start segment, and the unknown-tag error.  The reference's own tests only
call GM (src/c_geometry_tb.cpp:58-60) without checking coordinates."""
import numpy as np
import pytest

from oracle import pyoracle as po

KPI = 3.1415926536
# cos(90 deg) with the reference's pi is ~5e-12, not 0
EPS = 1e-10


def wire_x(ns=4, length=0.4, tag=1):
    g = po.Geometry()
    g.wire(tag, ns, 0.0, 0.0, 0.0, length, 0.0, 0.0, 0.001)
    return g


def centres(g):
    tab, _, _, tags = g.segments()
    return tab[:, 0:3], tab[:, 5:8], tags


def rot_z(deg):
    a = deg * KPI / 180.0
    return np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]])


def test_rotation_by_right_angles_and_translation():
    g = wire_x()
    g.gm(0, 0, 0.0, 0.0, 90.0, 1.0, 2.0, 3.0)  # about z, then shift
    g.complete(0)
    c, d, _ = centres(g)
    want = np.array([[1.0, 2.0 + 0.05 + 0.1 * k, 3.0] for k in range(4)])
    assert np.max(np.abs(c - want)) < EPS
    assert np.max(np.abs(d - [0.0, 1.0, 0.0])) < EPS
    # about y: x goes to -z (right-handed rotation of the x axis about y)
    g = wire_x()
    g.gm(0, 0, 0.0, 90.0, 0.0, 0.0, 0.0, 0.0)
    g.complete(0)
    c, d, _ = centres(g)
    assert np.max(np.abs(c - [[0.0, 0.0, -(0.05 + 0.1 * k)] for k in range(4)])) < EPS
    # about x leaves an x-directed wire alone
    g = wire_x()
    g.gm(0, 0, 37.0, 0.0, 0.0, 0.0, 0.0, 0.0)
    g.complete(0)
    c, _, _ = centres(g)
    assert np.max(np.abs(c - [[0.05 + 0.1 * k, 0.0, 0.0] for k in range(4)])) < 1e-15


def test_order_of_rotations_is_x_then_y_then_z():
    # a y-directed wire rotated by 90 about x (-> z) and then 90 about y
    # (-> x); z first would leave it along y
    g = po.Geometry()
    g.wire(1, 2, 0.0, 0.0, 0.0, 0.0, 0.2, 0.0, 0.001)
    g.gm(0, 0, 90.0, 90.0, 0.0, 0.0, 0.0, 0.0)
    g.complete(0)
    _, d, _ = centres(g)
    assert np.max(np.abs(d - [1.0, 0.0, 0.0])) < EPS


def test_copies_rotate_the_previous_copy_and_raise_tags():
    g = wire_x(ns=3, length=0.3, tag=2)
    g.gm(5, 3, 0.0, 0.0, 30.0, 0.0, 0.0, 0.0)
    g.complete(0)
    c, _, tags = centres(g)
    assert list(tags) == [2] * 3 + [7] * 3 + [12] * 3 + [17] * 3
    base = np.array([[0.05 + 0.1 * k, 0.0, 0.0] for k in range(3)])
    for r in range(4):
        assert np.max(np.abs(c[3 * r:3 * r + 3] - base @ rot_z(30.0 * r).T)) < 1e-12, r


def test_its_starts_at_the_first_segment_of_the_tag():
    g = po.Geometry()
    g.wire(1, 2, 0.0, 0.0, 0.0, 0.2, 0.0, 0.0, 0.001)
    g.wire(0, 2, 0.0, 1.0, 0.0, 0.2, 1.0, 0.0, 0.001)  # tag 0 stays 0 in copies
    g.wire(3, 2, 0.0, 2.0, 0.0, 0.2, 2.0, 0.0, 0.001)
    g.gm(4, 1, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0)  # all segments, one copy up by 1
    g.gm(0, 0, 0.0, 0.0, 0.0, 5.0, 0.0, 0.0, 3)  # from tag 3's first segment on
    g.complete(0)
    c, _, tags = centres(g)
    assert list(tags) == [1, 1, 0, 0, 3, 3, 5, 5, 0, 0, 7, 7]
    # before tag 3: untouched; tag 3 and everything after it: shifted in x
    assert np.max(np.abs(c[:4, 0] - [0.05, 0.15, 0.05, 0.15])) < 1e-15
    assert np.max(np.abs(c[4:, 0] - 5.0 - np.tile([0.05, 0.15], 4))) < 1e-14
    assert np.max(np.abs(c[6:, 2] - 1.0)) < 1e-15


def test_unknown_tag_and_scale():
    g = wire_x()
    with pytest.raises(RuntimeError, match="NO SEGMENT HAS AN ITAG OF 9"):
        g.gm(0, 1, 0.0, 0.0, 45.0, 0.0, 0.0, 0.0, 9)
    g.gs(2.5)
    g.complete(0)
    tab, _, _, _ = g.segments()
    assert np.max(np.abs(tab[:, 0] - [0.125 + 0.25 * k for k in range(4)])) < 1e-14
    assert np.allclose(tab[:, 4], 0.0025, rtol=1e-15)
