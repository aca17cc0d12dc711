"""Run a NEC card deck on the oracle (oracle/necpp_oracle.c through
pyoracle), for the deck tests: test infrastructure, independent of the
product's deck reader (necpp_amd/csrc/deck.cpp).

It reads the deck as nec2++ does (src/misc.cpp:74-160: a line starting with
'#' or a blank is skipped, a quote starts a comment, the first two
characters are the mnemonic; fields split on blanks and commas, missing
fields are zero) and follows nec_context's card semantics for the cards the
GPU path runs (src/nec_context.cpp:406-955, 1031-1159):
  geometry  GW (GC for a tapered wire), GM, GS, GE
  data      FR, EX 0 (consecutive EX cards add sources), LD (consecutive LD
            cards accumulate), GN, EK, KH, XQ, RP (mode 0), NE, NH, TL / NT
            (consecutive cards accumulate; a port I2 of -1 on the first clears
            them), PT/PQ (printing only), EN
A solve (fill, LU, currents) runs when XQ, RP, NE or NH needs one and the
structure, frequency or excitation changed since the last; RP and NE after
a solve at one frequency reuse it.  Results, in the order the reference
adds them: `inputs` (per solve and frequency: the input impedances of the
sources), `patterns` (per RP and frequency: the oracle's pattern dict),
`near` (per NE or NH: the field at the product's points, see
near_points and near_field), and
`currents` and `freqs` beside `inputs` (the segment currents and the
frequency of each solve).
"""
import numpy as np

from oracle import pyoracle as po


def cards(text):
    out = []
    for n, line in enumerate(text.splitlines(), 1):
        if not line or line[0] in "# ":
            continue
        line = line.split("'")[0]
        if not line:
            continue
        out.append((line[:2].upper(), line[2:].replace(",", " ").split(), n))
    return out


def fields(tok, ni, nr):
    iv = [int(float(t)) if t.lstrip("+-").replace(".", "", 1).isdigit() else int(t) for t in tok[:ni]]
    iv += [0] * (ni - len(iv))
    rv = [float(t) for t in tok[ni:ni + nr]]
    rv += [0.0] * (nr - len(rv))
    return iv, rv


def near_points(card):
    """The field points of an NE card in metres, in nfpat's order
    (src/nec_context.cpp:5068-5188): z (or theta) outermost, then y (phi),
    then x (r); counts of 0 read as 1; each coordinate accumulated by
    repeated addition as the reference steps it."""
    near, nrx, nry, nrz = (max(1, v) if i else v for i, v in enumerate(card[:4]))
    xnr, ynr, znr, dxnr, dynr, dznr = card[4:]
    d2r = (2 * 3.1415926536) / 360.0
    pts = []
    zz = znr - dznr
    for _ in range(nrz):
        zz += dznr
        yy = ynr - dynr
        for _ in range(nry):
            yy += dynr
            xx = xnr - dxnr
            for _ in range(nrx):
                xx += dxnr
                if near:
                    th, ph = zz * d2r, yy * d2r
                    pts.append((xx * np.sin(th) * np.cos(ph), xx * np.sin(th) * np.sin(ph), xx * np.cos(th)))
                else:
                    pts.append((xx, yy, zz))
    return np.array(pts)


def run(text, nthreads=8):
    cs = cards(text)
    k = 0
    if k < len(cs) and cs[k][0] in ("CM", "CE"):
        while cs[k][0] == "CM":
            k += 1
        k += 1
    g = po.Geometry()
    gpflag = 0
    while k < len(cs):
        name, tok, _ = cs[k]
        k += 1
        if name == "GW":
            iv, rv = fields(tok, 2, 7)
            rad, rdel, rrad = rv[6], 1.0, 1.0
            if rad == 0.0:
                _, gv = fields(cs[k][1], 2, 7)
                k += 1
                rdel, rad = gv[0], gv[1]
                rrad = (gv[2] / gv[1]) ** (1.0 / (iv[1] - 1.0))
            g.wire(iv[0], iv[1], *rv[:6], rad, rdel, rrad)
        elif name == "GS":
            g.gs(fields(tok, 2, 7)[1][0])
        elif name == "GM":
            iv, rv = fields(tok, 2, 7)
            g.gm(iv[0], iv[1], *rv[:6], its=int(rv[6] + 0.5))
        elif name == "GE":
            gpflag = fields(tok, 2, 7)[0][0]
            g.complete(gpflag)
            break
        else:
            raise ValueError(f"geometry card {name} not handled")
    res = {"inputs": [], "currents": [], "freqs": [], "patterns": [], "near": [], "geometry": g}
    freq = [po.speed_of_light() / 1.0e6]  # nec_context's default frequency (src/nec_context.cpp:158)
    src, loads, nets = [], [], []
    last = None
    use_ek, rkh, ground = False, 1.0, (-1, 0.0, 0.0)
    solved = {}  # frequency -> run_voltage result of the current structure and excitation

    def solve(f, rp=None):
        g.set_ground(*ground)
        g.set_loads(loads)
        g.set_networks(nets)
        segs = [(g.segment_number(t, s), v) for t, s, v in src]
        r = g.run_voltage(f, segs, rkh=rkh, use_ek=use_ek, rp=rp, nthreads=nthreads)
        r["freq"] = f
        return r

    while k < len(cs):
        name, tok, line = cs[k]
        k += 1
        iv, rv = fields(tok, 4, 6)
        if name == "EN":
            break
        if name == "FR":
            n = max(1, iv[1])
            f0, df = rv[0], rv[1]
            freq = [f0]
            for _ in range(n - 1):
                freq.append(freq[-1] * df if iv[0] == 1 else freq[-1] + df)
            solved = {}
        elif name == "EX":
            if iv[0] != 0:
                raise ValueError("only EX 0")
            if last != "EX":
                src = []
            v = complex(rv[0], rv[1])
            src.append((iv[1], iv[2], v if abs(v) >= 1e-20 else 1.0))  # 0 V reads as 1 V (src/nec_context.cpp:587-590)
            solved = {}
        elif name == "LD":
            if iv[0] == -1:
                loads = []
            else:
                loads.append((iv[0], iv[1], iv[2], iv[3], rv[0], rv[1], rv[2]))
            solved = {}
        elif name in ("TL", "NT"):
            # src/nec_context.cpp:630-726
            if last not in ("TL", "NT"):
                nets = []
                if iv[1] == -1:
                    last = name
                    solved = {}
                    continue
            nets.append((name, g.segment_number(iv[0], iv[1]), g.segment_number(iv[2], iv[3]), tuple(rv)))
            solved = {}
        elif name == "GN":
            ground = (iv[0], rv[0], rv[1]) if iv[0] != -1 else (-1, 0.0, 0.0)
            solved = {}
        elif name == "EK":
            use_ek = iv[0] != -1
            solved = {}
        elif name == "KH":
            rkh = rv[0]
            solved = {}
        elif name in ("PT", "PQ"):
            pass
        elif name == "XQ":
            if iv[0] != 0:
                raise ValueError("only XQ 0")
            if not (len(freq) == 1 and freq[0] in solved):
                for f in freq:
                    solved[f] = solve(f)
                    res["inputs"].append(solved[f]["zin"])
                    res["currents"].append(solved[f]["current"])
                    res["freqs"].append(f)
        elif name == "RP":
            if iv[0] != 0:
                raise ValueError("only RP mode 0")
            x = iv[3]
            rp = dict(n_theta=max(1, iv[1]), n_phi=max(1, iv[2]), output_format=x // 1000,
                      normalization=(x // 100) % 10, ipd=(x // 10) % 10, iavp=x % 10, theta0=rv[0], phi0=rv[1],
                      dtheta=rv[2], dphi=rv[3], rfld=rv[4], gnor=rv[5])
            for f in freq:
                if f in solved and len(freq) == 1:
                    pat = solve(f, rp)  # the same solve again, for its pattern: no new input result
                else:
                    pat = solve(f, rp)
                    solved[f] = pat
                    res["inputs"].append(pat["zin"])
                    res["currents"].append(pat["current"])
                    res["freqs"].append(f)
                res["patterns"].append(pat)
        elif name in ("NE", "NH"):
            if len(freq) != 1:
                raise ValueError("NE / NH with one frequency only")
            f = freq[0]
            if f not in solved:
                solved[f] = solve(f)
                res["inputs"].append(solved[f]["zin"])
                res["currents"].append(solved[f]["current"])
                res["freqs"].append(f)
            res["near"].append({"card": iv + rv, "abc": solved[f]["abc"], "freq": f, "rkh": rkh, "use_ek": use_ek,
                                "ground": ground, "magnetic": name == "NH"})
        else:
            raise ValueError(f"data card {name} (line {line}) not handled")
        last = name
    return res


def near_field(res, entry, pts_m):
    """The oracle's near field (E for NE, H for NH) of a near entry at
    points in metres, over the ground of its solve."""
    g = res["geometry"]
    f = entry["freq"]
    g.set_ground(*entry["ground"])
    g.scale(f)
    lam = po.speed_of_light() / (f * 1e6)
    pts = np.asarray(pts_m) / lam
    if entry["magnetic"]:
        return g.near_h(entry["abc"], pts)
    return g.near_e(entry["abc"], pts, rkh=entry["rkh"], use_ek=entry["use_ek"])
