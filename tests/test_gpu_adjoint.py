"""GPU: adjoint gradients (nec_gpu_adjoint_gradient) against central finite
differences of the objective the library itself reports.

J = -G(theta, phi) + weight * |Z_in - z0|^2 (G: total power gain in dBi of
the RP card, Z_in: nec_impedance_*).  The finite differences use the
relative step 1e-6 of the differentiation workload and evaluate J through
the unchanged C API (RP card gain and impedance), so they also check that
the J the adjoint reports is the J of the solve.
"""
import os

import numpy as np
import pytest

from necpp_amd import adjoint as adj
from necpp_amd import decks
from necpp_amd import necpp as api

pytestmark = pytest.mark.gpu

REL_STEP = 1e-6
# The junction structure's J is ~5.6e3 (a large impedance mismatch term) and
# its evaluations carry ~1e-13 relative noise, i.e. ~3e-3 absolute in a
# central difference at step 1e-7 m; the step 1e-5 keeps the differences'
# own error below 1e-7 relative there.
JUNCTION_STEP = 1e-5


# ----------------------------------------------------------------- a junction

def junction_deck(P, radii, freq_mhz, theta=60.0, phi=30.0):
    """Three wires meeting at junction point J: (A) from A0 to J, (B) from J
    to B1, (C) from J to C1.  P = [A0, J, B1, C1] (4 x 3 metres)."""
    A0, J, B1, C1 = (tuple(map(float, P[k])) for k in range(4))
    return {
        "wires": [(1, 7) + A0 + J + (radii[0], 1.0, 1.0),
                  (2, 5) + J + B1 + (radii[1], 1.0, 1.0),
                  (3, 6) + J + C1 + (radii[2], 1.0, 1.0)],
        "gpflag": 0,
        "freq_mhz": freq_mhz,
        "source": (1, 4),
        "voltage": 1.0 + 0.5j,
        "rp": (0, 1, 1, 1, 0, 0, 0, theta, phi, 0.0, 0.0, 0.0, 0.0),
    }


def deck_J(deck, weight, z0):
    ctx = api.nec_create()
    try:
        decks.apply_to_api(deck, api, ctx, with_rp=True)
        g = api.nec_gain(ctx, 0, 0, 0)
        z = complex(api.nec_impedance_real(ctx, 0), api.nec_impedance_imag(ctx, 0))
    finally:
        api.nec_delete(ctx)
    return -g + weight * abs(z - z0) ** 2


def deck_adjoint(deck, theta, phi, weight, z0):
    ctx = api.nec_create()
    try:
        decks.apply_to_api(deck, api, ctx, with_rp=False)
        return api.gpu_adjoint_gradient(ctx, theta, phi, weight, z0)
    finally:
        api.nec_delete(ctx)


def wire_end_gradient(g7, seg0, ns):
    """dJ/d(GW card endpoints) of a straight uniform wire from its segments'
    endpoint gradient: segment k spans P1 + k (P2 - P1) / ns .. P1 + (k + 1)
    (P2 - P1) / ns."""
    g = g7[seg0:seg0 + ns]
    k = np.arange(ns)[:, None]
    d1 = (g[:, 0:3] * (1 - k / ns) + g[:, 3:6] * (1 - (k + 1) / ns)).sum(axis=0)
    d2 = (g[:, 0:3] * (k / ns) + g[:, 3:6] * ((k + 1) / ns)).sum(axis=0)
    return d1, d2


def test_adjoint_junction_endpoints_radii_frequency():
    theta, phi, weight, z0 = 60.0, 30.0, 0.02, 60.0 + 10.0j
    P = np.array([[0.0, 0.0, -0.3], [0.0, 0.01, 0.0], [0.2, 0.0, 0.2], [-0.15, 0.1, 0.25]])
    radii = np.array([0.002, 0.003, 0.0025])
    f = 310.0
    J, g7, dfreq, info = deck_adjoint(junction_deck(P, radii, f, theta, phi), theta, phi, weight, z0)
    assert np.all(np.isfinite(g7))
    Jr = deck_J(junction_deck(P, radii, f, theta, phi), weight, z0)
    assert abs(J - Jr) <= 1e-12 * max(1.0, abs(Jr)), (J, Jr)
    # analytic gradients of the 12 point coordinates, radii and frequency
    A = wire_end_gradient(g7, 0, 7)
    B = wire_end_gradient(g7, 7, 5)
    C = wire_end_gradient(g7, 12, 6)
    grad_P = np.array([A[0], A[1] + B[0] + C[0], B[1], C[1]])
    grad_r = np.array([g7[0:7, 6].sum(), g7[7:12, 6].sum(), g7[12:18, 6].sum()])

    def fd(make):
        lo, hi, h = make()
        return (deck_J(hi, weight, z0) - deck_J(lo, weight, z0)) / (2 * h)

    worst = 0.0
    for q in range(4):
        for c in range(3):
            h = JUNCTION_STEP * max(abs(P[q, c]), 0.1)

            def make(q=q, c=c, h=h):
                Pp, Pm = P.copy(), P.copy()
                Pp[q, c] += h
                Pm[q, c] -= h
                return (junction_deck(Pm, radii, f, theta, phi), junction_deck(Pp, radii, f, theta, phi), h)
            num = fd(make)
            err = abs(grad_P[q, c] - num) / max(abs(num), 1e-3 * np.abs(grad_P).max())
            worst = max(worst, err)
            assert err < 1e-6, ("point", q, c, grad_P[q, c], num)
    for w in range(3):
        h = JUNCTION_STEP * radii[w]

        def make(w=w, h=h):
            rp, rm = radii.copy(), radii.copy()
            rp[w] += h
            rm[w] -= h
            return (junction_deck(P, rm, f, theta, phi), junction_deck(P, rp, f, theta, phi), h)
        num = fd(make)
        assert abs(grad_r[w] - num) <= 1e-6 * abs(num), ("radius", w, grad_r[w], num)
    h = JUNCTION_STEP * f
    num = fd(lambda: (junction_deck(P, radii, f - h, theta, phi), junction_deck(P, radii, f + h, theta, phi), h))
    assert abs(dfreq - num) <= 1e-6 * abs(num), ("frequency", dfreq, num)
    print("junction: worst point-coordinate relative error", worst)


# ---------------------------------------------------- the differentiation workload

FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "c4_fd_oracle.npz")


def fd_compare(G, fd, fd2):
    """Normwise relative error against fd, and the differences' own
    consistency |fd - fd2| / |fd| (steps h and 2h)."""
    nf = np.linalg.norm(fd)
    return np.linalg.norm(G - fd) / nf, np.linalg.norm(fd - fd2) / nf


def test_adjoint_all_designs_match_oracle_central_differences():
    """All 256 designs: the adjoint gradient against the oracle's central
    differences (tests/golden/make_c4_fd.py), every design to 1e-6 normwise.

    The differences hold the fill's geometry-dependent decisions fixed (the
    near/far switch at r = rkh, the adaptive quadrature's steps, recorded at p
    and replayed at p +- h), so they differentiate the same function the
    adjoint does; without that, a difference straddling a decision flip is
    off by the jump / 2h.  The bar is against the fourth-order difference of
    the step h = 1e-6 |p| (fd4 = (4 fd(h) - fd(2h)) / 3): for the radii that
    step is 5e-4 of the parameter, and fd(h)'s own h^2 error reaches 1e-6."""
    z = np.load(FIXTURE)
    assert bool(z["frozen"])
    P = decks.c4_params(count=len(z["J"]))
    assert np.array_equal(P.sum(axis=1), z["params_sum"])
    errs, errs_h = [], []
    for d, p in enumerate(P):
        J, G, _ = adj.objective_and_gradient(p)
        assert abs(J - z["J"][d]) <= 1e-10 * abs(z["J"][d]), (d, J, z["J"][d])
        nf = np.linalg.norm(z["fd4"][d])
        errs.append(np.linalg.norm(G - z["fd4"][d]) / nf)
        errs_h.append(np.linalg.norm(G - z["fd"][d]) / nf)
    errs, errs_h = np.array(errs), np.array(errs_h)
    print("designs", len(errs), "within 1e-6 of fd4", int((errs <= 1e-6).sum()), "median", np.median(errs),
          "max", errs.max(), "| of fd(h): within 1e-6", int((errs_h <= 1e-6).sum()), "max", errs_h.max())
    bad = np.nonzero(errs > 1e-6)[0]
    assert bad.size == 0, [(int(d), errs[d]) for d in bad]


def test_adjoint_yagi_design_matches_gpu_central_differences():
    # the same check with differences of the GPU path's own J (RP gain and
    # impedance through the C API) for the first design
    p = decks.c4_params(count=1)[0]
    J, G, info = adj.objective_and_gradient(p)
    assert abs(J - adj.objective(p)) <= 1e-12 * abs(J)
    num = adj.central_differences(p, REL_STEP)
    num2 = adj.central_differences(p, 2 * REL_STEP)
    e, c = fd_compare(G, num, num2)
    print("design 0: J", J, "error", e, "fd consistency", c, "info", info)
    assert c <= 1e-6 and e <= 1e-6


def test_adjoint_torch_function():
    import torch
    P = torch.tensor(decks.c4_params(count=2), dtype=torch.float64, requires_grad=True)
    J = adj.yagi_objective(P)
    J.sum().backward()
    Jb, Gb, _ = adj.batch_objective_and_gradient(P.detach().numpy())
    assert np.array_equal(J.detach().numpy(), Jb)
    assert np.array_equal(P.grad.numpy(), Gb)
    # a single row keeps its shape
    p = torch.tensor(decks.c4_params(count=1)[0], dtype=torch.float64, requires_grad=True)
    j = adj.yagi_objective(p)
    assert j.ndim == 0
    (2.0 * j).backward()
    assert np.array_equal(p.grad.numpy(), 2.0 * adj.batch_objective_and_gradient(p.detach().numpy()[None])[1][0])


def test_batch_adjoint_matches_single_path():
    """The batched adjoint (one call for all designs) against the
    single-structure adjoint: the batched LU rounds differently, so J to
    1e-12 and the gradient to 1e-9 normwise."""
    P = decks.c4_params(count=4)
    J, G, stats = adj.batch_objective_and_gradient(P)
    print("batch adjoint stats", stats)
    assert stats["solves"] == 4
    for d, p in enumerate(P):
        Js, Gs, _ = adj.objective_and_gradient(p)
        assert abs(J[d] - Js) <= 1e-12 * abs(Js)
        assert np.linalg.norm(G[d] - Gs) <= 1e-9 * np.linalg.norm(Gs), d


def test_batch_adjoint_junction_two_frequencies():
    theta, phi, weight, z0 = 60.0, 30.0, 0.02, 60.0 + 10.0j
    P = np.array([[0.0, 0.0, -0.3], [0.0, 0.01, 0.0], [0.2, 0.0, 0.2], [-0.15, 0.1, 0.25]])
    radii = np.array([0.002, 0.003, 0.0025])
    deck = junction_deck(P, radii, 310.0, theta, phi)
    freqs = [300.0, 320.0]
    J, grads, df, info, stats = api.gpu_batch_adjoint([deck], freqs, theta, phi, weight, z0)
    for f, fm in enumerate(freqs):
        Js, g7, dfs, _ = deck_adjoint(dict(deck, freq_mhz=fm), theta, phi, weight, z0)
        assert abs(J[0, f] - Js) <= 1e-12 * abs(Js)
        assert np.linalg.norm(grads[f] - g7) <= 1e-9 * np.linalg.norm(g7)
        assert abs(df[0, f] - dfs) <= 1e-9 * abs(dfs)


def test_batch_adjoint_all_designs_match_oracle_central_differences():
    """The 256 designs in one batched call, held to the same bar as the
    single-structure test above: every design within 1e-6 normwise."""
    z = np.load(FIXTURE)
    P = decks.c4_params(count=len(z["J"]))
    J, G, stats = adj.batch_objective_and_gradient(P)
    print("batch adjoint of 256 designs", stats)
    assert np.all(np.abs(J - z["J"]) <= 1e-10 * np.abs(z["J"]))
    errs = np.linalg.norm(G - z["fd4"], axis=1) / np.linalg.norm(z["fd4"], axis=1)
    bad = np.nonzero(errs > 1e-6)[0]
    assert bad.size == 0, [(int(d), errs[d]) for d in bad]


def test_adjoint_errors():
    deck = decks.c4_deck(decks.c4_params(count=1)[0])
    ctx = api.nec_create()
    try:
        with pytest.raises(RuntimeError, match="no solve"):
            api.gpu_adjoint_gradient(ctx)
        api.nec_gpu_set_precision(ctx, 1)
        decks.apply_to_api(deck, api, ctx, with_rp=False)
        with pytest.raises(RuntimeError, match="FP64"):
            api.gpu_adjoint_gradient(ctx)
    finally:
        api.nec_delete(ctx)
