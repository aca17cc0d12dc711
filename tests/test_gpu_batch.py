"""GPU: batched solves (nec_gpu_batch_solve) of Yagi populations.

Each (design, frequency) of a batch must give what the single-structure
path gives for the same deck (which tests/test_gpu_parity.py holds to the
oracle): Z_in within 1e-10 relative and the E-plane gain within 1e-9 dB.
The captured graph's replays must equal the eager pass bit for bit.
"""
import numpy as np
import pytest

from necpp_amd import decks
from necpp_amd import necpp as api
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

TOL = 1e-10


def single(deck, f, npoints=361):
    """Z_in and E-plane gains of one deck at f MHz through the C API."""
    ctx = api.nec_create()
    try:
        d = dict(deck)
        d["freq_mhz"] = f
        decks.apply_to_api(d, api, ctx, with_rp=True)
        z = complex(api.nec_impedance_real(ctx, 0), api.nec_impedance_imag(ctx, 0))
        g = api.gpu_pattern_array(ctx, 0, "pg_tot", npoints)
        return z, g
    finally:
        api.nec_delete(ctx)


def test_batch_matches_single_path_and_oracle():
    designs, _, theta, phi = decks.yagi_population(count=6, seed=7)
    freqs = np.array([140.0, 145.0, 150.0])
    zin, gain, stats = api.gpu_batch_solve(designs, freqs, theta, phi, repeats=2)
    print("batch stats", stats)
    assert stats["solves"] == 18
    assert stats["graph_matches_eager"] == 1
    for d, deck in enumerate(designs):
        for f, fm in enumerate(freqs):
            z, g = single(deck, fm)
            assert abs(zin[d, f] - z) <= TOL * abs(z), (d, f, zin[d, f], z)
            assert np.max(np.abs(gain[d, f] - g)) < 1e-9, (d, f)
    # and two of them against the oracle directly
    for d, f in ((0, 0), (5, 2)):
        o = po.run_deck(dict(designs[d], freq_mhz=freqs[f]), nthreads=4, with_rp=False)
        assert abs(zin[d, f] - o["zin"][0]) <= TOL * abs(o["zin"][0])


def test_batch_panel_widths_agree():
    designs, _, theta, phi = decks.yagi_population(count=5, seed=11)
    freqs = np.array([143.0, 147.0])
    ref = None
    for p in (8, 16, 32):
        zin, gain, stats = api.gpu_batch_solve(designs, freqs, theta[:10], phi[:10], panel_width=p)
        assert stats["panel_width"] == p
        if ref is None:
            ref = (zin, gain)
        else:
            assert np.max(np.abs(zin - ref[0]) / np.abs(ref[0])) < 1e-12
            assert np.max(np.abs(gain - ref[1])) < 1e-10


def test_batch_without_pattern_and_errors():
    designs, freqs, _, _ = decks.yagi_population(count=3, seed=3, n_freq=2)
    zin, gain, stats = api.gpu_batch_solve(designs, freqs)
    assert zin.shape == (3, 2) and gain.shape == (3, 2, 0)
    assert stats["panel_width"] in (8, 16, 32)
    big = decks.yagi([0.5] * 25, [0.3] * 24, 0.003, 21)  # 525 segments
    with pytest.raises(RuntimeError, match="batched path takes"):
        api.gpu_batch_solve([big], freqs)


def test_batch_large_structures_use_512_thread_workgroups():
    # 273 and 441 segments: the 512-thread kernels, against the single path
    rng = np.random.default_rng(4)
    lam0 = decks.C_NEC / 145.0e6
    designs = [decks.yagi(list(rng.uniform(0.2, 0.26, k) * lam0), list(rng.uniform(0.1, 0.3, k - 1) * lam0),
                          0.003 * lam0, 21) for k in (13, 21)]
    freqs = np.array([144.0, 146.0])
    theta, phi = np.full(5, 90.0), np.arange(5) * 45.0
    for p in (8, 16):
        zin, gain, stats = api.gpu_batch_solve(designs, freqs, theta, phi, panel_width=p)
        assert stats["panel_width"] == p
        for d, deck in enumerate(designs):
            for f, fm in enumerate(freqs):
                z, g = single(dict(deck, rp=(0, 1, 5, 1, 0, 0, 0, 90.0, 0.0, 0.0, 45.0, 0.0, 0.0)), fm, 5)
                assert abs(zin[d, f] - z) <= TOL * abs(z), (p, d, f)
                assert np.max(np.abs(gain[d, f] - g)) < 1e-9


def test_batch_full_population_sample():
    # the full workload: 512 designs x 64 frequencies, 361-point cuts
    designs, freqs, theta, phi = decks.yagi_population()
    zin, gain, stats = api.gpu_batch_solve(designs, freqs, theta, phi, repeats=1)
    print("C3 stats", stats)
    assert stats["solves"] == 512 * 64
    assert stats["graph_matches_eager"] == 1
    assert np.all(np.isfinite(zin)) and np.all(np.isfinite(gain))
    rng = np.random.default_rng(5)
    for d, f in zip(rng.integers(0, 512, 6), rng.integers(0, 64, 6)):
        z, g = single(designs[d], freqs[f])
        assert abs(zin[d, f] - z) <= TOL * abs(z), (d, f)
        assert np.max(np.abs(gain[d, f] - g)) < 1e-9, (d, f)
    # a stratified sample straight against the oracle: the first design of
    # every element count (3 .. 12) at a random frequency of the sweep
    counts = [len(dk["wires"]) for dk in designs]
    rp = dict(n_theta=1, n_phi=361, theta0=90.0, phi0=0.0, dtheta=0.0, dphi=1.0)
    for k in range(3, 13):
        d = counts.index(k)
        f = int(rng.integers(0, 64))
        dk = dict(designs[d])
        dk["freq_mhz"] = float(freqs[f])
        g = po.geometry_from_deck(dk)
        o = g.run_voltage(dk["freq_mhz"], [(g.segment_number(*dk["source"]), dk["voltage"])], rp=rp, nthreads=8)
        assert abs(zin[d, f] - o["zin"][0]) <= TOL * abs(o["zin"][0]), (k, d, f)
        og = o["pattern"]["pg_tot"]
        assert np.max(np.abs(gain[d, f] - og)) < 1e-8, (k, d, f)
