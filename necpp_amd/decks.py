"""Benchmark and test structures as lists of GW cards.

Each deck is a dict with
  wires:  list of (tag, ns, x1, y1, z1, x2, y2, z2, rad, rdel, rrad)
  gpflag: GE flag
  freq_mhz, source (tag, segment_in_tag), voltage
and optionally rp (the nec_rp_card arguments).  `apply(deck, sink)` feeds
the cards to anything with nec_wire / nec_geometry_complete style methods.
"""

C_REF = 299792458.0  # only used to size the Yagi designs in wavelengths


def dipole(ns=51, half=0.25, rad=0.001, freq_mhz=299.792458, rp=(181, 361)):
    """Centre-fed straight dipole along z (the N=51 CPU-runnable case)."""
    deck = {
        "name": f"dipole{ns}",
        "wires": [(1, ns, 0.0, 0.0, -half, 0.0, 0.0, half, rad, 1.0, 1.0)],
        "gpflag": 0,
        "freq_mhz": freq_mhz,
        "source": (1, ns // 2 + 1),
        "voltage": 1.0 + 0.0j,
    }
    if rp:
        deck["rp"] = (0, rp[0], rp[1], 1, 0, 0, 0, 0.0, 0.0, 1.0, 1.0, 0.0, 0.0)
    return deck


def turnstile(phase=-1j, ns=21, half=0.25, sep=0.05, freq_mhz=299.792458):
    """Two crossed half-wave dipoles, x-directed at z = 0 and y-directed at
    z = sep, fed at their centre segments with 1 V and `phase` V (two
    sources: "sources" lists ((tag, segment), voltage)).  Fed in quadrature
    the far field is elliptically polarised, nearly circular along z."""
    return {
        "name": "turnstile",
        "wires": [(1, ns, -half, 0.0, 0.0, half, 0.0, 0.0, 0.001, 1.0, 1.0),
                  (2, ns, 0.0, -half, sep, 0.0, half, sep, 0.001, 1.0, 1.0)],
        "gpflag": 0,
        "freq_mhz": freq_mhz,
        "sources": [((1, ns // 2 + 1), 1.0 + 0.0j), ((2, ns // 2 + 1), complex(phase))],
    }


def _node(k, nn, spacing):
    return float(f"{-(nn - 1) * spacing / 2.0 + k * spacing:.10f}")


def grid(nn=41, spacing=0.05, rad=0.001, freq_mhz=299.792458, rp=(91, 361)):
    """Planar square wire grid in z = 0 with nn x nn nodes, one segment per edge.

    Horizontal edges row by row first, then vertical edges column by column;
    tag k+1 for the k-th edge.  The source is the first edge whose centre is
    nearest the origin (nn=41 -> tag 820, N=3,280; nn=90 -> tag 3961,
    N=16,020).
    """
    xs = [_node(k, nn, spacing) for k in range(nn)]
    wires = []
    for r in range(nn):
        for c in range(nn - 1):
            wires.append((0, 1, xs[c], xs[r], 0.0, xs[c + 1], xs[r], 0.0, rad, 1.0, 1.0))
    for c in range(nn):
        for r in range(nn - 1):
            wires.append((0, 1, xs[c], xs[r], 0.0, xs[c], xs[r + 1], 0.0, rad, 1.0, 1.0))
    best, btag = None, None
    out = []
    for k, w in enumerate(wires):
        tag = k + 1
        out.append((tag,) + w[1:])
        cx, cy = 0.5 * (w[2] + w[5]), 0.5 * (w[3] + w[6])
        d = cx * cx + cy * cy
        if best is None or d < best - 1e-15:
            best, btag = d, tag
    deck = {
        "name": f"grid{nn}",
        "wires": out,
        "gpflag": 0,
        "freq_mhz": freq_mhz,
        "source": (btag, 1),
        "voltage": 1.0 + 0.0j,
    }
    if rp:
        deck["rp"] = (0, rp[0], rp[1], 1, 0, 0, 0, 0.0, 0.0, 1.0, 1.0, 0.0, 0.0)
    return deck


def yagi(half_lengths, spacings, rad, seg_per_element=21, freq_mhz=145.0, rp=(1, 361)):
    """Yagi-Uda array: elements parallel to y, boom along x, source on the
    centre segment of element 2 (lengths in metres)."""
    wires = []
    x = 0.0
    for e, h in enumerate(half_lengths):
        if e > 0:
            x += spacings[e - 1]
        wires.append((e + 1, seg_per_element, x, -h, 0.0, x, h, 0.0, rad, 1.0, 1.0))
    deck = {
        "name": f"yagi{len(half_lengths)}x{seg_per_element}",
        "wires": wires,
        "gpflag": 0,
        "freq_mhz": freq_mhz,
        "source": (2, seg_per_element // 2 + 1),
        "voltage": 1.0 + 0.0j,
    }
    if rp:
        # E-plane cut: theta = 90 deg (the array plane), phi 0..360
        deck["rp"] = (0, rp[0], rp[1], 1, 0, 0, 0, 90.0, 0.0, 0.0, 1.0, 0.0, 0.0)
    return deck


# speed of light of nec2++ (electromag.cpp: 1 / sqrt(mu0 eps0), eps0 = 8.854e-12,
# mu0 = 4 pi 1e-7 with pi = 3.1415926536)
C_NEC = 1.0 / (4.0 * 3.1415926536e-7 * 8.854e-12) ** 0.5
BATCH_SEED = 20261203


def yagi_population(count=512, seed=BATCH_SEED, n_freq=64, f_lo=140.0, f_hi=150.0, seg_per_element=21):
    """The batching workload: `count` Yagi-Uda designs drawn from numpy's
    PCG64 generator seeded with `seed`.  Per design, in draw order: the
    element count k ~ U{3..12}, k half-lengths ~ U(0.20, 0.28) lambda0 and
    k-1 spacings ~ U(0.10, 0.35) lambda0 (lambda0 at 145 MHz); radius 0.003
    lambda0; source on the centre segment of element 2.  Each design is swept
    over n_freq frequencies evenly spaced over [f_lo, f_hi] MHz and cut in
    its E-plane (theta = 90 deg, phi = 0..360 deg in 1 deg steps).
    Returns (decks, freqs, theta_deg, phi_deg)."""
    import numpy as np
    lam0 = C_NEC / 145.0e6
    rng = np.random.default_rng(seed)
    designs = []
    for _ in range(count):
        k = int(rng.integers(3, 13))
        half = rng.uniform(0.20, 0.28, k) * lam0
        spac = rng.uniform(0.10, 0.35, k - 1) * lam0
        designs.append(yagi(list(half), list(spac), 0.003 * lam0, seg_per_element, 145.0))
    freqs = np.linspace(f_lo, f_hi, n_freq)
    phi = np.arange(361, dtype=float)
    theta = np.full(361, 90.0)
    return designs, freqs, theta, phi


ADJOINT_SEED = 20261204
ADJOINT_FREQ_MHZ = 144.0


def c4_lambda(freq_mhz=ADJOINT_FREQ_MHZ):
    """Wavelength (m) of nec2++ at freq_mhz."""
    return C_NEC / (freq_mhz * 1.0e6)


def c4_params(count=256, seed=ADJOINT_SEED, n_elem=15):
    """The differentiation workload's design parameters: `count` rows of
    3 n_elem - 1 values in wavelengths, drawn from numpy's PCG64 generator
    seeded with `seed`, per design in order: n_elem half-lengths ~ U(0.20,
    0.26), n_elem - 1 spacings ~ U(0.10, 0.30), n_elem radii ~ U(0.002,
    0.005)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    rows = []
    for _ in range(count):
        h = rng.uniform(0.20, 0.26, n_elem)
        d = rng.uniform(0.10, 0.30, n_elem - 1)
        r = rng.uniform(0.002, 0.005, n_elem)
        rows.append(np.concatenate([h, d, r]))
    return np.array(rows)


def c4_deck(p, freq_mhz=ADJOINT_FREQ_MHZ, seg_per_element=31):
    """Yagi-Uda deck of one parameter row (c4_params): elements parallel to
    y, boom along x, element k at x = sum of the first k spacings, source on
    the centre segment of element 2; pattern cut: the single direction
    theta = 90, phi = 0 (forward, +x)."""
    lam = c4_lambda(freq_mhz)
    ne = (len(p) + 1) // 3
    h, d, r = p[:ne], p[ne:2 * ne - 1], p[2 * ne - 1:]
    wires = []
    x = 0.0
    for e in range(ne):
        if e > 0:
            x += d[e - 1] * lam
        wires.append((e + 1, seg_per_element, x, -h[e] * lam, 0.0, x, h[e] * lam, 0.0, r[e] * lam, 1.0, 1.0))
    return {
        "name": f"yagi{ne}x{seg_per_element}-adjoint",
        "wires": wires,
        "gpflag": 0,
        "freq_mhz": freq_mhz,
        "source": (2, seg_per_element // 2 + 1),
        "voltage": 1.0 + 0.0j,
        "rp": (0, 1, 1, 1, 0, 0, 0, 90.0, 0.0, 0.0, 0.0, 0.0, 0.0),
    }


def apply_to_api(deck, api, ctx, with_rp=True):
    """Run the deck through a libnecpp-style API object (module or wrapper)."""
    for w in deck["wires"]:
        rc = api.nec_wire(ctx, *w)
        if rc != 0:
            raise RuntimeError(api.nec_error_message())
    for call, args in (("nec_geometry_complete", (deck["gpflag"],)),
                       ("nec_fr_card", (0, 1, deck["freq_mhz"], 0.0)),
                       ("nec_ex_card", (0, deck["source"][0], deck["source"][1], 0,
                                        deck["voltage"].real, deck["voltage"].imag, 0.0, 0.0, 0.0, 0.0))):
        rc = getattr(api, call)(ctx, *args)
        if rc != 0:
            raise RuntimeError(api.nec_error_message())
    if with_rp and "rp" in deck:
        rc = api.nec_rp_card(ctx, *deck["rp"])
    else:
        rc = api.nec_xq_card(ctx, 0)
    if rc != 0:
        raise RuntimeError(api.nec_error_message())
