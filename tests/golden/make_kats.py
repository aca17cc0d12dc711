"""Write tests/golden/kats.json: the known-answer values that pin this path.

Sources (read as text, nothing is imported or executed):
  * nec2++'s own Python tests, testharness/python/test_examples.py: the
    assertAlmostEqual literals of test_example1 (free-space 7-segment dipole,
    fill -> LU -> solve -> Z_in) are parsed out of the file.
  * nec2++'s C++ unit tests: src/nec_context_tb.cpp:47-48 (the same dipole,
    tolerance 3e-4) and src/matrix_algebra_tb.cpp (LU / solve known answers,
    tolerance 1e-4): transcribed below with their line numbers.
  * Values measured through the reference's C API (nec_impedance_real/imag,
    nec_gain) on the five benchmark decks, recorded with their provenance in
    the repository's baseline notes; transcribed below.

Run: python tests/golden/make_kats.py [/path/to/nec2++/tree]
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def parse_example1(ref_root):
    path = os.path.join(ref_root, "testharness", "python", "test_examples.py")
    text = open(path).read()
    m = re.search(r"def test_example1\(self\):(.*?)def test_example2", text, re.S)
    body = m.group(1)
    re_v = re.search(r"nec_impedance_real\(nec,0\),([0-9.eE+-]+)\)", body).group(1)
    im_v = re.search(r"nec_impedance_imag\(nec,0\),([0-9.eE+-]+)\)", body).group(1)
    return float(re_v), float(im_v)


def main():
    ref_root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    zr, zi = parse_example1(ref_root)
    kats = {
        "example1_dipole_7seg": {
            "source": "testharness/python/test_examples.py:27-36 (assertAlmostEqual, 7 places)",
            "deck": {"wire": [0, 7, 0.0, 0.0, 0.75, 0.0, 0.0, 1.25, 0.001, 1.0, 1.0], "gpflag": 1,
                     "freq_mhz": "default (c/1e6)", "ex": [0, 0, 4, 0, 1.0, 0.0]},
            "zin": [zr, zi],
            "zin_reference_now": [82.697929066626244, 46.306038880634347],
            "zin_reference_now_source": "reference C API probe, recorded in the survey (8c)",
        },
        "nec_context_tb_example1": {
            "source": "src/nec_context_tb.cpp:14-48 (REQUIRE |d| < 3e-4)",
            "deck": {"wire": [0, 7, 0.0, 0.0, -0.25, 0.0, 0.0, 0.25, 0.001, 1.0, 1.0], "gpflag": 0,
                     "ex": [0, 0, 4, 0, 1.0, 0.0]},
            "zin": [82.698, 46.306],
            "tol_abs": 3e-4,
        },
        "c1_dipole_51seg": {
            "source": "reference C API probe (nec_impedance_real/imag, nec_gain, nec_gain_max) on the "
                      "N=51 benchmark deck; recorded in BASELINE.md section 2",
            "deck": {"wire": [1, 51, 0.0, 0.0, -0.25, 0.0, 0.0, 0.25, 0.001, 1.0, 1.0], "gpflag": 0,
                     "fr": [0, 1, 299.792458, 0.0], "ex": [0, 1, 26, 0, 1.0, 0.0],
                     "rp": [0, 181, 361, 1, 0, 0, 0, 0.0, 0.0, 1.0, 1.0, 0.0, 0.0]},
            "zin": [85.966843405368309, 48.88324221285977],
            "gain_theta90_phi0": 2.1826018307326764,
            "gain_max": 2.1826018307326764,
        },
        "c2_grid_3280": {
            "source": "reference CLI print (5 significant digits), BASELINE.md section 2",
            "deck": "41x41-node grid, 0.05 m spacing, see tests/decks.py",
            "source_tag": 820,
            "zin": [1.1838, 146.10],
            "tol_rel_print": 5e-4,
        },
        "c5_grid_16020": {
            "source": "reference CLI print (5 significant digits), BASELINE.md section 2",
            "deck": "90x90-node grid, 0.05 m spacing, see tests/decks.py",
            "source_tag": 3961,
            "zin": [1.3646, 146.20],
            "tol_rel_print": 5e-4,
        },
        "lu_4x4": {
            "source": "src/matrix_algebra_tb.cpp:87-129 (x = i+1 - j(i+1)/2, |d| < 1e-4)",
            "A": [[3.0, 1.0, -4.0, 2.0], [3.0, 1.0, 0.0, 2.0], [2.0, 13.0, -1.0, 0.0], [-2.0, 3.0, -1.0, 4.0]],
        },
        "lu_12x12": {
            "source": "src/matrix_algebra_tb.cpp:164-221 (diag 3(i+1)+j(i+1), off 0.1(i-j)+j0.05(i+1)(j+1)/N^2)",
            "n": 12,
        },
    }
    with open(os.path.join(HERE, "kats.json"), "w") as f:
        json.dump(kats, f, indent=1)
    print("wrote", os.path.join(HERE, "kats.json"))


if __name__ == "__main__":
    main()
