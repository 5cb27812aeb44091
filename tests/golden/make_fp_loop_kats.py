"""Write tests/golden/fp_loop_kats.json: the known answers of the reference's
arbitrary-precision loop tests, read as text from tests/test_FP_loop.py of the
Fractalshades source tree; nothing of the reference is executed.

* ``ref_path``: Test_ref_path (test_FP_loop.py:88-158, 201-212): the orbit of
  the image centre (no Newton) of a frame at 3520 digits, dx = 2e-2608,
  max_iter = 10^7; the iterations where |Z| < 1e-300 (the Xrange register)
  are the 14 listed in test_numba_path's docstring, and the membership checks
  of numba_path_loop (test_FP_loop.py:66-80).
* ``ball_method``: Test_newton.test_ball_method (test_FP_loop.py:166-193).
* ``newton``: Test_newton.test_newton (test_FP_loop.py:231-272): the
  divisor-removal Newton of z^2 + c from a start point at order 4252,
  11622 bits, 80 steps; the nucleus to ~3500 digits.
* ``bs_path``: Test_BS.test_path (test_FP_loop.py:278-309): 10 steps of the
  burning ship from c = 0.1 + 0.1i at 53 bits.
* ``bs_order``: Test_BS.test_order (test_FP_loop.py:311-341): ball-method
  period 3 at two points, px 0.01, 100 steps, M 1000.
* ``bs_newton``: Test_BS.test_newton (test_FP_loop.py:343-382): the 2x2
  Newton at order 3 from the same two points (px 0.01 and 0.012, 40 steps,
  eps_cv 2^-53); the reference asserts only that it runs.
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

    nw = txt[txt.index("def test_newton(self):"):txt.index("BURNING_SHIP = 1")]
    newton = {k: re.search(r'"' + k + r'":\s*"([-0-9.e]+)"', nw).group(1)
              for k in ("x_start", "y_start", "x_nucleus", "y_nucleus")}
    newton["order"] = int(re.search(r'"order":\s*(\d+)', nw).group(1))
    d, f = re.search(r"precision = int\((\d+) \* ([0-9.]+)\)", nw).groups()
    newton["precision_bits"] = int(int(d) * float(f))
    d, f = re.search(r"pix = int\((\d+) \* ([0-9.]+)\)", nw).groups()
    newton["pix_bits"] = int(int(d) * float(f))
    newton["max_newton"] = int(re.search(r'newton_search\["order"\],\s*(\d+),', nw).group(1))
    newton["source"] = "tests/test_FP_loop.py:231-272"

    bs = txt[txt.index("class Test_BS"):]
    tp = bs[bs.index("def test_path"):bs.index("def test_order")]
    bs_path = {"x": float(re.search(r"\bx = ([-0-9.e]+)", tp).group(1)),
               "y": float(re.search(r"\by = ([-0-9.e]+)", tp).group(1)),
               "npts": int(re.search(r"npts = (\d+)", tp).group(1)),
               "M": float(re.search(r"npts - 1,\s*([0-9.e]+),", tp).group(1)),
               "precision_bits": int(re.search(r"\),\s*(\d+),\s*BURNING_SHIP", tp).group(1)),
               "source": "tests/test_FP_loop.py:278-309"}
    to = bs[bs.index("def test_order"):bs.index("def test_newton")]
    pts = re.findall(r'x_str = "([-0-9.e]+)"\s*y_str = "([-0-9.e]+)"\s*seed_px = "([0-9.e]+)"', to)
    orders = [int(v) for v in re.findall(r"assert order == (\d+)", to)]
    bs_order = {"points": [list(p) for p in pts], "orders": orders,
                "precision_bits": 53, "maxiter": 100, "M_divergence": 1000.,
                "source": "tests/test_FP_loop.py:311-341"}
    tn = bs[bs.index("def test_newton"):]
    pts = re.findall(r'x_str = "([-0-9.e]+)"\s*y_str = "([-0-9.e]+)"\s*seed_px = "([0-9.e]+)"\s*'
                     r'seed_prec = (\d+)\s*order = (\d+)', tn.split("#")[0])
    bs_newton = {"cases": [{"x": p[0], "y": p[1], "px": p[2], "prec": int(p[3]), "order": int(p[4])}
                           for p in pts],
                 "max_newton": int(re.search(r"order,\s*(\d+),", tn).group(1)),
                 "source": "tests/test_FP_loop.py:343-382"}

    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fp_loop_kats.json")
    with open(path, "w") as fh:
        json.dump({"ref_path": ref, "ball_method": ball, "newton": newton, "bs_path": bs_path,
                   "bs_order": bs_order, "bs_newton": bs_newton}, fh, indent=1)


if __name__ == "__main__":
    main()
