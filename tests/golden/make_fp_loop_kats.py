"""Write tests/golden/fp_loop_kats.json: the known answers of the reference's
arbitrary-precision loop tests, read as text from tests/test_FP_loop.py of the
Fractalshades source tree; nothing of the reference is executed.

* ``ref_path``: Test_ref_path (test_FP_loop.py:88-158, 201-212): the orbit of
  the image centre (no Newton) of a frame at 3520 digits, dx = 2e-2608,
  max_iter = 10^7; the iterations where |Z| < 1e-300 (the Xrange register)
  are the 14 listed in test_numba_path's docstring, and the membership checks
  of numba_path_loop (test_FP_loop.py:66-80).
* ``ball_method``: Test_newton.test_ball_method (test_FP_loop.py:166-193).
"""
import json
import os
import re

SRC = "/root/reference/tests/test_FP_loop.py"


def main():
    txt = open(SRC).read()
    s = txt[txt.index("class Test_ref_path"):]
    ref = {
        "x": re.search(r"\bx = '([-0-9.e]+)'", s).group(1),
        "y": re.search(r"\by = '([-0-9.e]+)'", s).group(1),
        "dx": re.search(r"\bdx = '([-0-9.e]+)'", s).group(1),
        "precision": int(re.search(r"precision = (\d+)", s).group(1)),
        "nx": int(re.search(r"\bnx = (\d+)", s).group(1)),
        "max_iter": int(re.search(r"max_iter=(\d+)", s).group(1)),
        "M_divergence": float(re.search(r"M_divergence=([0-9.e]+)", s).group(1)),
    }
    doc = re.search(r"14 xr indices\s*\[([0-9\s]+)\]", s).group(1)
    ref["xr_indices"] = [int(v) for v in doc.split()]
    loop = txt[txt.index("def numba_path_loop"):txt.index("def numba_c_from_pix")]
    idx = re.search(r"for i in \(([0-9, ]+)\):", loop).group(1)
    yn = re.search(r"#\s+([YN ]+)\n", loop).group(1).split()
    ref["membership"] = [[int(i), f == "Y"] for i, f in zip(idx.split(","), yn)]
    ref["source"] = "tests/test_FP_loop.py:66-80, 88-158, 201-212"

    b = txt[txt.index("def test_ball_method"):]
    ball = {k: re.search(r'"' + k + r'":\s*"?([-0-9.e]+)"?', b).group(1)
            for k in ("x", "y", "maxiter", "radius", "M_divergence", "order")}
    ball["maxiter"] = int(ball["maxiter"])
    ball["order"] = int(ball["order"])
    ball["M_divergence"] = float(ball["M_divergence"])
    ball["precision_bits"] = int(78 * 3.33)       # test_FP_loop.py:181
    ball["source"] = "tests/test_FP_loop.py:166-193"

    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fp_loop_kats.json")
    with open(path, "w") as fh:
        json.dump({"ref_path": ref, "ball_method": ball}, fh, indent=1)


if __name__ == "__main__":
    main()
