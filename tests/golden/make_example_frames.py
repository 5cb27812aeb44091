"""Write tests/golden/example_frames.json: zoom locations (x, y, dx, skew,
max_iter) of the reference's burning-ship batch examples, read as text from
examples/batch_mode/ of the Fractalshades source tree.  Used as parity
frames; nothing of the reference is executed."""
import json
import os
import re

SRC = "/root/reference/examples/batch_mode"
FILES = {
    # name: (file, flavor, decimal precision when the file does not set one)
    "BS_ex14": ("14-burning_ship-deeper_DEM.py", "Burning ship", None),
    "PBS_ex18": ("18-perpendicular_burning_ship_glynn.py", "Perpendicular burning ship", 55),
    "SF_ex22": ("22-shark_fin_deep.py", "Shark fin", 725),
}


def grab(txt, key):
    m = re.search(r"^\s*" + key + r"\s*=\s*['\"]?([^'\",\s]+)['\"]?", txt, re.M)
    return m.group(1) if m else None


def main():
    out = {}
    for name, (fn, flavor, prec) in FILES.items():
        txt = open(os.path.join(SRC, fn)).read()
        ent = {k: grab(txt, k) for k in ("x", "y", "dx", "precision", "max_iter",
                                          "skew_00", "skew_01", "skew_10", "skew_11")}
        ent["precision"] = int(ent["precision"]) if ent["precision"] else prec
        ent["max_iter"] = int(ent["max_iter"])
        for k in ("skew_00", "skew_01", "skew_10", "skew_11"):
            ent[k] = float(ent[k])
        ent["flavor"] = flavor
        ent["source"] = f"examples/batch_mode/{fn}"
        out[name] = ent
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "example_frames.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)


if __name__ == "__main__":
    main()
