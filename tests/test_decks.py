"""CPU: the benchmark workloads are the ones the configurations describe."""
import numpy as np

from necpp_amd import decks


def test_yagi_population_is_seeded_and_in_range():
    a, fa, th, ph = decks.yagi_population(count=40)
    b, fb, _, _ = decks.yagi_population(count=40)
    assert [d["wires"] for d in a] == [d["wires"] for d in b]
    assert np.array_equal(fa, fb)
    lam0 = decks.C_NEC / 145.0e6
    for d in a:
        k = len(d["wires"])
        assert 3 <= k <= 12
        n = sum(w[1] for w in d["wires"])
        assert n == 21 * k and 63 <= n <= 252
        half = np.array([w[6] for w in d["wires"]]) / lam0
        assert np.all((half >= 0.20) & (half <= 0.28))
        xs = np.array([w[2] for w in d["wires"]])
        sp = np.diff(xs) / lam0
        assert np.all((sp >= 0.10 - 1e-12) & (sp <= 0.35 + 1e-12))
        assert all(abs(w[8] - 0.003 * lam0) < 1e-15 for w in d["wires"])
        assert d["source"] == (2, 11)
    assert len(fa) == 64 and fa[0] == 140.0 and fa[-1] == 150.0
    assert np.all(th == 90.0) and np.array_equal(ph, np.arange(361.0))


def test_nec_speed_of_light():
    # 1 / sqrt(mu0 eps0) with the reference's truncated pi
    assert abs(decks.C_NEC - 299795637.69) < 0.01
