"""Pin the CPU oracle against the reference's known answers.

The oracle (oracle/necpp_oracle.c) is the checker for every GPU parity test,
so it is itself checked here against nec2++'s own known-answer tests and the
values measured through the reference's C API (tests/golden/kats.json).
"""
import os

import numpy as np
import pytest

from necpp_amd import decks
from oracle import pyoracle as po


def rel(a, b):
    return abs(a - b) / abs(b)


def test_example1_kat(kats):
    # testharness/python/test_examples.py:27-36: 16-digit Z_in of a
    # free-space 7-segment dipole (fill -> LU -> solve -> Z_in).
    k = kats["example1_dipole_7seg"]
    g = po.Geometry()
    g.wire(*k["deck"]["wire"])
    g.complete(k["deck"]["gpflag"])
    r = g.run_voltage(po.speed_of_light() / 1.0e6, [(4, 1.0)])
    z = r["zin"][0]
    ref = complex(*k["zin"])
    assert rel(z, ref) < 1e-13
    assert rel(z, complex(*k["zin_reference_now"])) < 1e-13
    # the reference test's own criterion (7 places)
    assert round(z.real - ref.real, 7) == 0 and round(z.imag - ref.imag, 7) == 0


def test_nec_context_tb_example1(kats):
    # src/nec_context_tb.cpp:14-48
    k = kats["nec_context_tb_example1"]
    g = po.Geometry()
    g.wire(*k["deck"]["wire"])
    g.complete(0)
    z = g.run_voltage(po.speed_of_light() / 1.0e6, [(4, 1.0)])["zin"][0]
    assert abs(z.real - k["zin"][0]) < k["tol_abs"]
    assert abs(z.imag - k["zin"][1]) < k["tol_abs"]


def test_c1_dipole_kat(kats):
    # N = 51 dipole: Z_in and gain values measured through the reference's C API
    k = kats["c1_dipole_51seg"]
    r = po.run_deck(decks.dipole(), nthreads=4)
    assert rel(r["zin"][0], complex(*k["zin"])) < 1e-12
    gain = r["pattern"]["gain"]
    assert abs(gain[0 * 181 + 90] - k["gain_theta90_phi0"]) < 1e-11
    assert abs(r["max_gain"] - k["gain_max"]) < 1e-11
    assert abs(r["stats"][0] - k["gain_max"]) < 1e-11


def test_lu_4x4_known_solution(kats):
    # src/matrix_algebra_tb.cpp:87-129
    A = np.array(kats["lu_4x4"]["A"], dtype=np.complex128)
    x = np.array([(i + 1) - 0.5j * (i + 1) for i in range(4)])
    b = A @ x
    LU, piv, fz = po.lu(A)
    assert fz == -1
    xs = po.lu_solve(LU, piv, b)
    assert np.max(np.abs(xs - x)) < 1e-12


def test_lu_12x12_known_solution():
    # src/matrix_algebra_tb.cpp:164-221
    N = 12
    A = np.zeros((N, N), np.complex128)
    for i in range(N):
        for j in range(N):
            if i == j:
                A[i, j] = complex(3.0 * (i + 1), 1.0 * (i + 1))
            else:
                A[i, j] = complex(0.1 * (i - j), 0.05 * ((i + 1) * (j + 1)) / (N * N))
    x = np.array([(i + 1) - 0.5j * (i + 1) for i in range(N)])
    b = A @ x
    LU, piv, _ = po.lu(A)
    assert np.max(np.abs(po.lu_solve(LU, piv, b) - x)) < 1e-12


def test_lu_plu_reconstruction_and_pivot_rule():
    # P A = L U, and the pivot of every column is the first maximum of |z|
    rng = np.random.default_rng(7)
    n = 70
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    A[5] = A[9]  # a tie in magnitude in column 0 of rows 5 and 9 below
    LU, piv, _ = po.lu(A)
    L = np.tril(LU, -1) + np.eye(n)
    U = np.triu(LU)
    PA = A.copy()
    for k in range(n):
        PA[[k, piv[k]]] = PA[[piv[k], k]]
    assert np.max(np.abs(PA - L @ U)) < 1e-12 * np.max(np.abs(A))
    # replay the elimination and check every pivot choice
    W = A.copy()
    for k in range(n):
        col = np.abs(W[k:, k])
        assert piv[k] == k + int(np.argmax(col))
        W[[k, piv[k]]] = W[[piv[k], k]]
        W[k + 1:, k] /= W[k, k]
        W[k + 1:, k + 1:] -= np.outer(W[k + 1:, k], W[k, k + 1:])


def test_lu_matches_numpy_on_random():
    rng = np.random.default_rng(3)
    for n in (1, 2, 17, 64, 65, 130):
        A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) + n * np.eye(n)
        b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        LU, piv, _ = po.lu(A, nthreads=2)
        assert np.allclose(po.lu_solve(LU, piv, b), np.linalg.solve(A, b), rtol=1e-12, atol=1e-13)


@pytest.mark.slow
def test_c2_grid_kat(kats):
    # N = 3,280 grid: Z_in against the reference's printed value (5 digits)
    k = kats["c2_grid_3280"]
    d = decks.grid(41, rp=None)
    assert d["source"] == (k["source_tag"], 1)
    r = po.run_deck(d, nthreads=8, with_rp=False)
    z = r["zin"][0]
    assert abs(z.real - k["zin"][0]) < 0.51e-4 * 10
    assert abs(z.imag - k["zin"][1]) < 0.51e-2 * 10
    assert rel(z, complex(*k["zin"])) < k["tol_rel_print"]


def test_c5_fixture_matches_reference_print(kats):
    # The N = 16,020 oracle fixture (tests/golden/make_c5_oracle.py) against
    # the reference's printed Z_in (5 significant digits); the deck's source
    # is the reference's tag 3961.
    import json
    k = kats["c5_grid_16020"]
    f = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "c5_oracle.json")))
    d = decks.grid(90, rp=None)
    assert d["source"] == (k["source_tag"], 1)
    assert f["n"] == len(d["wires"]) == 16020
    z = complex(*f["zin"])
    assert rel(z, complex(*k["zin"])) < k["tol_rel_print"]
    cur = np.load(os.path.join(os.path.dirname(__file__), "golden", "c5_oracle_current.npy"))
    assert cur.shape == (16020,)
    assert abs(np.linalg.norm(cur) - f["current_l2"]) < 1e-15
    # Z_in = V / I at the source segment (1 V source; segment numbers 1-based)
    assert rel(z, 1.0 / cur[f["source_segment_number"] - 1]) < 1e-12
    assert rel(complex(*f["current_src"]), cur[f["source_segment_number"] - 1]) == 0
