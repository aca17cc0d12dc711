"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Every test runs the product library (libnecpp_amd.so) on the GPU and checks
it against oracle/necpp_oracle.c on the same inputs, and against the
reference's known answers where they exist.  Bar: FP64 results within 1e-10
relative; the fill keeps the reference's
operation order, so it is held to 1e-12.
"""
import ctypes
import os

import numpy as np
import pytest
import torch  # noqa: F401  (imported at collection: the first import on a fresh box is slow)

from necpp_amd import decks
from necpp_amd import necpp as api
from oracle import pyoracle as po
from gpu_compare import check_stats, compare_pattern

pytestmark = pytest.mark.gpu

TOL = 1e-10


def make(deck, ek=False):
    ctx = api.nec_create()
    assert ctx
    for w in deck["wires"]:
        assert api.nec_wire(ctx, *w) == 0
    assert api.nec_geometry_complete(ctx, deck["gpflag"]) == 0, api.nec_error_message()
    if ek:
        assert api.nec_ek_card(ctx, 0) == 0
    return ctx


def run(deck, with_rp=True, ek=False):
    ctx = make(deck, ek)
    assert api.nec_fr_card(ctx, 0, 1, deck["freq_mhz"], 0.0) == 0
    s = deck["source"]
    v = deck["voltage"]
    assert api.nec_ex_card(ctx, 0, s[0], s[1], 0, v.real, v.imag, 0, 0, 0, 0) == 0, api.nec_error_message()
    if with_rp and "rp" in deck:
        assert api.nec_rp_card(ctx, *deck["rp"]) == 0, api.nec_error_message()
    else:
        assert api.nec_xq_card(ctx, 0) == 0, api.nec_error_message()
    return ctx


def zin(ctx, k=0):
    return complex(api.nec_impedance_real(ctx, k), api.nec_impedance_imag(ctx, k))


def rel(a, b):
    return abs(a - b) / abs(b)


# ---------------------------------------------------------------- LU / solve

def _torch():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a visible MI355X"
    return torch


def test_mfma_f64_lane_layout():
    # The LU's trailing update assumes, for v_mfma_f64_16x16x4f64:
    # A[i][k] from lane i + 16k, B[k][j] from lane j + 16k, and result
    # register v of lane l holding D[l/16 + 4*v][l % 16].
    torch = _torch()
    rng = np.random.default_rng(11)
    A = rng.standard_normal((16, 4))
    B = rng.standard_normal((4, 16))
    lanes = np.arange(64)
    a = torch.from_numpy(A[lanes % 16, lanes // 16].copy()).to("cuda")
    b = torch.from_numpy(B[lanes // 16, lanes % 16].copy()).to("cuda")
    d = torch.zeros(256, dtype=torch.float64, device="cuda")
    assert api.necgpu_check_mfma_f64(a.data_ptr(), b.data_ptr(), d.data_ptr(), None) == 0
    torch.cuda.synchronize()
    got = d.cpu().numpy().reshape(64, 4)
    D = A @ B
    found = {}
    for l in range(64):
        for v in range(4):
            hit = np.argwhere(np.abs(D - got[l, v]) < 1e-12)
            found[(l, v)] = tuple(hit[0]) if len(hit) == 1 else None
    print("MFMA f64 16x16x4 layout (lane, reg) -> (i, j):", {k: found[k] for k in list(found)[:20]})
    expect = {(l, v): (l // 16 + 4 * v, l % 16) for l in range(64) for v in range(4)}
    assert found == expect


def test_mfma_f32_lane_layout():
    # The mixed-precision trailing update assumes, for v_mfma_f32_16x16x4f32:
    # A[i][k] from lane i + 16k, B[k][j] from lane j + 16k, and result
    # register v of lane l holding D[4*(l/16) + v][l % 16].
    torch = _torch()
    rng = np.random.default_rng(12)
    A = rng.standard_normal((16, 4)).astype(np.float32)
    B = rng.standard_normal((4, 16)).astype(np.float32)
    lanes = np.arange(64)
    a = torch.from_numpy(A[lanes % 16, lanes // 16].copy()).to("cuda")
    b = torch.from_numpy(B[lanes // 16, lanes % 16].copy()).to("cuda")
    d = torch.zeros(256, dtype=torch.float32, device="cuda")
    assert api.necgpu_check_mfma_f32(a.data_ptr(), b.data_ptr(), d.data_ptr(), None) == 0
    torch.cuda.synchronize()
    got = d.cpu().numpy().reshape(64, 4).astype(np.float64)
    D = A.astype(np.float64) @ B.astype(np.float64)
    found = {}
    for l in range(64):
        for v in range(4):
            hit = np.argwhere(np.abs(D - got[l, v]) < 1e-5)
            found[(l, v)] = tuple(int(x) for x in hit[0]) if len(hit) == 1 else None
    print("MFMA f32 16x16x4 layout (lane, reg) -> (i, j):", {k: found[k] for k in list(found)[:20]})
    expect = {(l, v): (4 * (l // 16) + v, l % 16) for l in range(64) for v in range(4)}
    assert found == expect


# NECGPU_ZGEMM selects the FP64 trailing-update kernel: "1" the persistent
# tile kernel everywhere (small n never reaches the one-stream phase that
# uses it by default), "0" the two-per-CU kernel, "2" the default split
@pytest.mark.parametrize("mode", ["2", "1", "0"])
@pytest.mark.parametrize("n", [1, 2, 5, 63, 64, 65, 127, 130, 257, 300, 641])
def test_zgetrf_zgetrs_random(n, mode, monkeypatch):
    monkeypatch.setenv("NECGPU_ZGEMM", mode)
    torch = _torch()
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    dA = torch.from_numpy(A.copy()).to("cuda")
    piv = torch.zeros(2 * n, dtype=torch.int32, device="cuda")
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    lw = torch.empty(int(api.necgpu_zgetrf_workspace(n)), dtype=torch.uint8, device="cuda")
    assert api.necgpu_zgetrf(dA.data_ptr(), n, n, piv.data_ptr(), info.data_ptr(), lw.data_ptr(), None) == 0
    torch.cuda.synchronize()
    assert int(info.item()) == 0
    LU = dA.cpu().numpy()
    # oracle factorisation of the same matrix: same pivots, same factors
    oLU, opiv, _ = po.lu(A)
    p = piv.cpu().numpy()
    assert np.array_equal(p[:n], opiv)
    scale = np.max(np.abs(oLU))
    assert np.max(np.abs(LU - oLU)) <= 1e-12 * scale * max(1, n / 64)
    # the composed permutation agrees with the transpositions
    perm = np.arange(n)
    for k in range(n):
        perm[[k, opiv[k]]] = perm[[opiv[k], k]]
    assert np.array_equal(p[n:], perm)
    ws = torch.zeros(int(api.necgpu_zgetrs_workspace(n)), dtype=torch.uint8, device="cuda")
    for epoch in (1, 2):
        x = torch.from_numpy(b.copy()).to("cuda")
        assert api.necgpu_zgetrs(dA.data_ptr(), n, n, piv.data_ptr(), x.data_ptr(), ws.data_ptr(), epoch, None) == 0
        torch.cuda.synchronize()
        xs = x.cpu().numpy()
        ref = np.linalg.solve(A, b)
        assert np.linalg.norm(xs - ref) <= TOL * np.linalg.norm(ref)
    assert np.linalg.norm(A @ xs - b) <= 1e-12 * np.linalg.norm(b) * n


@pytest.mark.parametrize("pairs", ["1", "0"])
@pytest.mark.parametrize("n", [129, 130, 191, 257, 300, 641, 1100])
def test_zgetrf_one_stream_panel_pairs(n, pairs, monkeypatch):
    """Without the look-ahead every panel runs on one stream, in pairs with
    one 128-deep trailing update per pair (the default) or one update per
    panel: the oracle's pivots and factors either way."""
    monkeypatch.setenv("NECGPU_LU_LOOKAHEAD", "0")
    monkeypatch.setenv("NECGPU_LU_PAIRS", pairs)
    torch = _torch()
    rng = np.random.default_rng(3 * n + 1)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    dA = torch.from_numpy(A.copy()).to("cuda")
    piv = torch.zeros(2 * n, dtype=torch.int32, device="cuda")
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    lw = torch.empty(int(api.necgpu_zgetrf_workspace(n)), dtype=torch.uint8, device="cuda")
    assert api.necgpu_zgetrf(dA.data_ptr(), n, n, piv.data_ptr(), info.data_ptr(), lw.data_ptr(), None) == 0
    torch.cuda.synchronize()
    assert int(info.item()) == 0
    oLU, opiv, _ = po.lu(A)
    assert np.array_equal(piv.cpu().numpy()[:n], opiv)
    scale = np.max(np.abs(oLU))
    assert np.max(np.abs(dA.cpu().numpy() - oLU)) <= 1e-12 * scale * max(1, n / 64)


@pytest.mark.parametrize("n", [70, 300])
def test_zgetrs_backward_stable_when_nearly_singular(n):
    """A matrix with a cluster of nearly dependent columns (1e-13 apart) has
    diagonal blocks of U with condition numbers near 1e13: the solve must
    still leave the residual of substitution, ||A x - b|| <= n u ||A|| ||x||
    (u = 2^-53), as the oracle's solve does on the same factors."""
    torch = _torch()
    rng = np.random.default_rng(7 + n)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    for k in range(n - 6, n):  # the last block: a rank-deficient cluster
        A[:, k] = A[:, n - 7] + 1e-13 * (rng.standard_normal(n) + 1j * rng.standard_normal(n))
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    dA = torch.from_numpy(A.copy()).to("cuda")
    piv = torch.zeros(2 * n, dtype=torch.int32, device="cuda")
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    lw = torch.empty(int(api.necgpu_zgetrf_workspace(n)), dtype=torch.uint8, device="cuda")
    assert api.necgpu_zgetrf(dA.data_ptr(), n, n, piv.data_ptr(), info.data_ptr(), lw.data_ptr(), None) == 0
    ws = torch.zeros(int(api.necgpu_zgetrs_workspace(n)), dtype=torch.uint8, device="cuda")
    x = torch.from_numpy(b.copy()).to("cuda")
    assert api.necgpu_zgetrs(dA.data_ptr(), n, n, piv.data_ptr(), x.data_ptr(), ws.data_ptr(), 1, None) == 0
    torch.cuda.synchronize()
    xs = x.cpu().numpy()
    u = 2.0 ** -53
    scale = np.linalg.norm(A, 2) * np.linalg.norm(xs)
    assert np.linalg.norm(xs) > 1e8 * np.linalg.norm(b) / np.linalg.norm(A, 2)  # the system is nearly singular
    assert np.linalg.norm(A @ xs - b) <= n * u * scale, np.linalg.norm(A @ xs - b) / scale
    # the same factors solved by the oracle (substitution) for comparison
    LU = dA.cpu().numpy()
    xo = po.lu_solve(LU, piv.cpu().numpy()[:n].astype(np.int32), b)
    print("backward error gpu", np.linalg.norm(A @ xs - b) / scale,
          "oracle", np.linalg.norm(A @ xo - b) / (np.linalg.norm(A, 2) * np.linalg.norm(xo)))


def test_zgetrf_reports_zero_pivot():
    torch = _torch()
    n = 70
    A = np.zeros((n, n), np.complex128)
    A[:, :3] = np.random.default_rng(1).standard_normal((n, 3))
    dA = torch.from_numpy(A).to("cuda")
    piv = torch.zeros(2 * n, dtype=torch.int32, device="cuda")
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    lw = torch.empty(int(api.necgpu_zgetrf_workspace(n)), dtype=torch.uint8, device="cuda")
    assert api.necgpu_zgetrf(dA.data_ptr(), n, n, piv.data_ptr(), info.data_ptr(), lw.data_ptr(), None) == 0
    torch.cuda.synchronize()
    assert int(info.item()) == 4  # 1 + first exactly-zero pivot (Eigen's first_zero_pivot)


# ---------------------------------------------------------------- fill

FILL_DECKS = [
    ("dipole7", decks.dipole(7, rp=None), False),
    ("dipole51", decks.dipole(51, rp=None), False),
    ("grid9", decks.grid(9, rp=None), False),
    ("yagi4", decks.yagi([0.5, 0.48, 0.45, 0.44], [0.4, 0.3, 0.35], 0.003, 11, rp=None), False),
    ("dipole51_ek", decks.dipole(51, rp=None), True),
    ("yagi4_ek", decks.yagi([0.5, 0.48, 0.45, 0.44], [0.4, 0.3, 0.35], 0.003, 11, rp=None), True),
]


@pytest.mark.parametrize("name,deck,ek", FILL_DECKS, ids=[d[0] for d in FILL_DECKS])
def test_fill_matches_oracle(name, deck, ek):
    ctx = make(deck, ek)
    try:
        Z = api.gpu_fill_matrix(ctx, deck["freq_mhz"])
        g = po.geometry_from_deck(deck)
        g.scale(deck["freq_mhz"])
        oZ, counts = g.fill(use_ek=ek, nthreads=4)
        assert counts["near_pairs"] > 0
        err = np.abs(Z - oZ)
        bound = 1e-12 * np.abs(oZ) + 1e-15 * np.max(np.abs(oZ))
        assert np.all(err <= bound), (np.max(err / (np.abs(oZ) + 1e-300)), np.count_nonzero(err > bound))
    finally:
        api.nec_delete(ctx)


# a sloping wire and a horizontal one joined above ground, one end grounded
BENT = [(1, 8, 0.0, 0.0, 0.0, 0.1, 0.0, 0.3, 0.002, 1.0, 1.0),
        (2, 9, 0.1, 0.0, 0.3, 0.45, 0.2, 0.3, 0.002, 1.0, 1.0)]
YAGI4 = decks.yagi([0.5, 0.48, 0.45, 0.44], [0.4, 0.3, 0.35], 0.003, 11, rp=None)
ONE_PASS_CASES = {
    "grid9": (decks.grid(9, rp=None), False, None, None),
    "grid15_ek": (decks.grid(15, rp=None), True, None, None),
    "yagi4_loads": (YAGI4, False, None, [(0, 2, 5, 7, 12.0, 2.0e-9, 8.0e-11), (4, 0, 0, 0, 3.0, 40.0, 0.0)]),
    "bent_finite_ground": ({"wires": BENT, "gpflag": 1, "freq_mhz": 299.792458}, False, (0, 0, 13.0, 0.005), None),
    "bent_perfect_ground_ek": ({"wires": BENT, "gpflag": 1, "freq_mhz": 299.792458}, True, (1, 0, 0.0, 0.0), None),
}


@pytest.mark.parametrize("case", sorted(ONE_PASS_CASES))
def test_one_pass_fill_equals_field_buffer_fill(case, monkeypatch):
    """The one-pass fill (fields of a cluster of bases in LDS) against the
    field-buffer fill (fields to HBM, then the gather), bit for bit, with
    clusters of 8 bases so that sources are shared between clusters; and
    both against the oracle's fill."""
    deck, ek, gn, loads = ONE_PASS_CASES[case]

    def fill(fused):
        monkeypatch.setenv("NECGPU_FILL_TILE", "8")
        monkeypatch.setenv("NECGPU_FILL_FUSED", "1" if fused else "0")
        ctx = make(deck, ek)
        try:
            if gn is not None:
                assert api.nec_gn_card(ctx, gn[0], gn[1], gn[2], gn[3], 0, 0, 0, 0) == 0, api.nec_error_message()
            for ld in loads or []:
                assert api.nec_ld_card(ctx, *ld) == 0, api.nec_error_message()
            return api.gpu_fill_matrix(ctx, deck["freq_mhz"])
        finally:
            api.nec_delete(ctx)

    Z1, Z0 = fill(True), fill(False)
    assert Z1.shape[0] > 8
    assert np.array_equal(Z1.view(np.float64), Z0.view(np.float64))
    g = po.geometry_from_deck(deck)
    if gn is not None:
        g.set_ground(gn[0], gn[2], gn[3])
    if loads:
        g.set_loads(loads)
    g.scale(deck["freq_mhz"])
    oZ, _ = g.fill(use_ek=ek, nthreads=4)
    err = np.abs(Z1 - oZ)
    bound = 1e-12 * np.abs(oZ) + 1e-15 * np.max(np.abs(oZ))
    assert np.all(err <= bound), (np.max(err / (np.abs(oZ) + 1e-300)), np.count_nonzero(err > bound))


# ---------------------------------------------------------------- end to end

def test_example1_kat_gpu(kats):
    k = kats["example1_dipole_7seg"]
    ctx = api.nec_create()
    try:
        assert api.nec_wire(ctx, *k["deck"]["wire"]) == 0
        assert api.nec_geometry_complete(ctx, 1) == 0
        assert api.nec_ex_card(ctx, 0, 0, 4, 0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0) == 0
        assert api.nec_xq_card(ctx, 0) == 0, api.nec_error_message()
        z = zin(ctx)
        assert rel(z, complex(*k["zin"])) < 1e-12
        assert rel(z, complex(*k["zin_reference_now"])) < 1e-12
    finally:
        api.nec_delete(ctx)


def test_c1_dipole_gpu(kats):
    k = kats["c1_dipole_51seg"]
    deck = decks.dipole()
    ctx = run(deck)
    try:
        o = po.run_deck(deck, nthreads=8)
        z = zin(ctx)
        assert rel(z, o["zin"][0]) < TOL
        assert rel(z, complex(*k["zin"])) < TOL
        cur = api.gpu_currents(ctx, 0)
        assert np.linalg.norm(cur - o["current"]) <= TOL * np.linalg.norm(o["current"])
        assert abs(api.nec_gain(ctx, 0, 90, 0) - k["gain_theta90_phi0"]) < 1e-9
        assert abs(api.nec_gain_max(ctx, 0) - k["gain_max"]) < 1e-9
        # all twelve pattern arrays and the twelve statistics
        g = compare_pattern(ctx, o, 181 * 361)
        check_stats(ctx, o, g)
    finally:
        api.nec_delete(ctx)


def test_c2_grid_gpu(kats):
    k = kats["c2_grid_3280"]
    deck = decks.grid(41)
    ctx = run(deck)
    try:
        z = zin(ctx)
        assert rel(z, complex(*k["zin"])) < k["tol_rel_print"]
        o = po.run_deck(deck, nthreads=16, with_rp=False)
        assert rel(z, o["zin"][0]) < TOL
        cur = api.gpu_currents(ctx, 0)
        assert np.linalg.norm(cur - o["current"]) <= TOL * np.linalg.norm(o["current"])
        t = api.gpu_stage_times(ctx)
        print("C2 stage times (ms):", t)
    finally:
        api.nec_delete(ctx)


@pytest.mark.parametrize("precision", [0, 1])
def test_timed_out_panel_exchange_refactors_on_one_stream(precision, monkeypatch):
    """A factorisation whose panel exchange timed out is repeated with every
    launch on one stream (the FP64 path refilling the matrix the first
    attempt overwrote) and gives the oracle's solution.  The timeout is
    forced from the host (NECGPU_TEST_PANEL_TIMEOUT): the kernels run
    normally."""
    deck = decks.grid(21, rp=None)
    o = po.run_deck(deck, nthreads=16, with_rp=False)
    monkeypatch.setenv("NECGPU_TEST_PANEL_TIMEOUT", "1")
    ctx = make(deck)
    try:
        assert api.nec_gpu_set_precision(ctx, precision) == 0
        s, v = deck["source"], deck["voltage"]
        assert api.nec_fr_card(ctx, 0, 1, deck["freq_mhz"], 0.0) == 0
        assert api.nec_ex_card(ctx, 0, s[0], s[1], 0, v.real, v.imag, 0, 0, 0, 0) == 0
        assert api.nec_xq_card(ctx, 0) == 0, api.nec_error_message()
        assert api.gpu_factor_fallbacks(ctx)[0] == 1
        assert rel(zin(ctx), o["zin"][0]) < TOL
        cur = api.gpu_currents(ctx, 0)
        assert np.linalg.norm(cur - o["current"]) <= TOL * np.linalg.norm(o["current"])
    finally:
        api.nec_delete(ctx)


def test_frequency_sweep_and_reuse():
    deck = decks.dipole(21, rp=(19, 5))
    ctx = make(deck)
    try:
        assert api.nec_fr_card(ctx, 0, 3, 250.0, 25.0) == 0
        assert api.nec_ex_card(ctx, 0, 1, 11, 0, 1.0, 0.0, 0, 0, 0, 0) == 0
        assert api.nec_rp_card(ctx, 0, 19, 5, 1, 0, 0, 1, 0.0, 0.0, 10.0, 90.0, 0, 0) == 0, api.nec_error_message()
        g = po.geometry_from_deck(deck)
        for m, f in enumerate((250.0, 275.0, 300.0)):
            o = g.run_voltage(f, [(11, 1.0)], rp=dict(n_theta=19, n_phi=5, theta0=0, phi0=0, dtheta=10,
                                                         dphi=90, output_format=1, iavp=1))
            assert rel(zin(ctx, m), o["zin"][0]) < TOL
            assert abs(api.nec_gain_max(ctx, m) - o["stats"][0]) < 1e-9
            assert abs(api.nec_gain(ctx, m, 9, 2) - o["pattern"]["gain"][2 * 19 + 9]) < 1e-9
        assert api.nec_impedance_real(ctx, 3) == -999.0
        # a new source at the last frequency re-uses the factors
        assert api.nec_ex_card(ctx, 0, 1, 5, 0, 0.5, 0.25, 0, 0, 0, 0) == 0
        assert api.nec_xq_card(ctx, 0) == 0, api.nec_error_message()
        o = g.run_voltage(300.0, [(5, 0.5 + 0.25j)])
        assert rel(zin(ctx, 3), o["zin"][0]) < TOL
    finally:
        api.nec_delete(ctx)


def test_two_feeds_indexed():
    deck = decks.dipole(31, rp=None)
    ctx = make(deck)
    try:
        assert api.nec_fr_card(ctx, 0, 1, 280.0, 0) == 0
        assert api.nec_ex_card(ctx, 0, 1, 10, 0, 1.0, 0.0, 0, 0, 0, 0) == 0
        assert api.nec_ex_card(ctx, 0, 1, 22, 0, 0.0, 1.0, 0, 0, 0, 0) == 0
        assert api.nec_xq_card(ctx, 0) == 0
        o = po.geometry_from_deck(deck).run_voltage(280.0, [(10, 1.0), (22, 1j)])
        for f in range(2):
            z = complex(api.nec_impedance_real_indexed(ctx, 0, f), api.nec_impedance_imag_indexed(ctx, 0, f))
            assert rel(z, o["zin"][f]) < TOL
    finally:
        api.nec_delete(ctx)


def test_device_allocation_failure_is_reported():
    deck = decks.dipole(51, rp=None)
    ctx = make(deck)
    os.environ["NECGPU_FAIL_ALLOC"] = str(40000)
    try:
        assert api.nec_xq_card(ctx, 0) == 1
        assert "out of memory" in api.nec_error_message()
    finally:
        del os.environ["NECGPU_FAIL_ALLOC"]
        api.nec_delete(ctx)


# the mixed path's FP32 LU on its own: P A = L U to FP32 rounding, for both
# pivot modes and every trailing-update kernel
@pytest.mark.parametrize("gemm", ["2", "1", "0"])
@pytest.mark.parametrize("pivot_mode", [1, 0])
@pytest.mark.parametrize("n", [1, 5, 64, 65, 130, 300, 641, 1100])
def test_cgetrf_factors(n, pivot_mode, gemm, monkeypatch):
    monkeypatch.setenv("NECGPU_ZGEMM", gemm)
    torch = _torch()
    rng = np.random.default_rng(1000 + n)
    A = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))).astype(np.complex64)
    dA = torch.from_numpy(A.copy()).to("cuda")
    piv = torch.zeros(2 * n, dtype=torch.int32, device="cuda")
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    lw = torch.empty(int(api.necgpu_zgetrf_workspace(n)), dtype=torch.uint8, device="cuda")
    assert api.necgpu_cgetrf(dA.data_ptr(), n, n, piv.data_ptr(), info.data_ptr(), lw.data_ptr(), pivot_mode,
                             None) == 0, api.nec_error_message()
    torch.cuda.synchronize()
    assert int(info.item()) == 0
    LU = dA.cpu().numpy().astype(np.complex128)
    p = piv.cpu().numpy()
    perm = np.arange(n)
    for k in range(n):
        perm[[k, p[k]]] = perm[[p[k], k]]
    assert np.array_equal(p[n:], perm)
    if pivot_mode == 1:
        # pivots stay inside the 64-row diagonal block of their column
        assert np.all(p[:n] // 64 == np.arange(n) // 64)
    L = np.tril(LU, -1) + np.eye(n)
    U = np.triu(LU)
    err = np.abs(L @ U - A[perm].astype(np.complex128))
    bound = 8 * n * 6e-8 * (np.abs(L) @ np.abs(U))
    assert np.all(err <= bound + 1e-30), float(np.max(err / (bound + 1e-30)))


def test_cgetrf_rejects_bad_pivot_mode():
    assert api.necgpu_cgetrf(None, 4, 4, None, None, None, 2, None) == 1
    assert "pivot_mode" in api.nec_error_message()
