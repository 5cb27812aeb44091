"""Write tests/golden/edge_frames.json: the zoom locations and calculation
options of the reference's edge-case frames, read as text from
tests/test_perturbation.py of the Fractalshades source tree (nothing of the
reference is executed):

* ``M2_E213``: test_M2_E213 (203-297), deep zoom with calc_orbit (field lines);
* ``M2_flake``: test_glitch_dyn (503-576), Dinkydau's "Flake", the dynamic
  glitch case the reference gates with its tightest image threshold (0.002);
* ``M2_ultradeep``: test_ultradeep_interior_detect (651-716), an Xrange frame
  (5.07e-433) with interior detection;
* ``M2_supersampling``: test_supersampling (376-420), 3x3 supersampling.

The calculation options follow the test's ``calc`` helper (717-760):
calc_fast frames run without interior detection and dz/dc, the others with
calc_orbit and backshift 2.
"""
import json
import os
import re

SRC = "/root/reference/tests/test_perturbation.py"

CASES = {
    "M2_E213": ("test_M2_E213", "203-297"),
    "M2_flake": ("test_glitch_dyn", "503-576"),
    "M2_ultradeep": ("test_ultradeep_interior_detect", "651-716"),
    "M2_supersampling": ("test_supersampling", "376-420"),
}


def body(txt, fn):
    b = txt[txt.index("def " + fn + "("):]
    return b[:b.index("\n    def ", 10)]


def main():
    txt = open(SRC).read()
    out = {}
    for name, (fn, lines) in CASES.items():
        b = body(txt, fn)
        g = lambda k: re.search(k + r"""\s*=\s*['"]([-0-9.e]+)['"]""", b).group(1)
        e = {"x": g(r"\bx"), "y": g(r"\by"), "dx": g(r"\bdx"),
             "precision": int(re.search(r"precision = (\d+)", b).group(1)),
             "nx": int(re.search(r"\bnx = (\d+)", b).group(1)),
             "source": f"tests/test_perturbation.py:{lines}"}
        m = re.search(r"max_iter\s*=\s*(\d+)", b)
        e["max_iter"] = int(m.group(1)) if m else 50000            # calc()'s default
        m = re.search(r"xy_ratio=([0-9./]+)\)", b)
        e["xy_ratio"] = eval(m.group(1), {}) if m else 1.0        # e.g. 16/9.
        e["calc_fast"] = "calc_fast=True" in b
        e["interior_detect"] = "interior_detect=True" in b
        m = re.search(r'supersampling="(\d+)x\1"', b)
        e["supersampling"] = int(m.group(1)) if m else None
        out[name] = e
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "edge_frames.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)


if __name__ == "__main__":
    main()
