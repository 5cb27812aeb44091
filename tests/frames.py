"""Test frames: the reference test suite's zoom locations
(tests/test_perturbation.py of Fractalshades) at reduced image sizes and
iteration caps, plus helpers to run them through the oracle."""
import json
import os

import numpy as np

import fractalshades_amd as fsa
from fractalshades_amd import configs

# name: (model, zoom kwargs, calc kwargs) -- locations from the reference
# tests (test_perturbation.py line numbers in the comments)
FRAMES = {
    # test_M2_E20, 66-69
    "M2_E20": ("M2", dict(x="-1.74928893611435556407228", y="0.", dx="5.e-20", precision=40),
               dict(max_iter=5000)),
    # test_M2_int_E11, 138-141
    "M2_E11": ("M2", dict(x="-1.74920463345912691e+00", y="-2.8684660237361114e-04",
                          dx="5e-12", precision=30), dict(max_iter=5000)),
    # test_glitch_divref, 308-312
    "M2_divref": ("M2", dict(x="-1.36768994867991128", y="0.00949048853859240532",
                             dx="2.477633848347765e-8", precision=40), dict(max_iter=2000)),
    # test_glitch_divref5, 431-435
    "M2_divref5": ("M2", dict(x="-1.8583883137588246496170054595",
                              y="0.000062628913856889509589212598191",
                              dx="4.688599337510725e-19", precision=50), dict(max_iter=5000)),
    # test_deep_interior_detect, 584-588 (interior detection)
    "M2_int32": ("M2", dict(x="-1.9072475829123560219806014933444037282313",
                            y="7.54605431570315771754553123674989535657416e-16",
                            dx="3.660343578223062e-32", precision=50),
                 dict(max_iter=20000, interior_detect=True)),
}


# burning-ship family: locations of the reference's batch examples
# (tests/golden/example_frames.json, written by make_example_frames.py)
_EX = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                  "golden", "example_frames.json")))
for _name, _e in _EX.items():
    FRAMES[_name] = ("BS", dict(x=_e["x"], y=_e["y"], dx=_e["dx"], precision=_e["precision"],
                               has_skew=True, skew_00=_e["skew_00"], skew_01=_e["skew_01"],
                               skew_10=_e["skew_10"], skew_11=_e["skew_11"]),
                     dict(max_iter=min(_e["max_iter"], 20000)))
    FRAMES[_name + "_flavor"] = _e["flavor"]

# the reference's edge cases (tests/golden/edge_frames.json, written by
# make_edge_frames.py): options of the test's calc helper
# (test_perturbation.py:717-760): calc_fast frames without dz/dc, the others
# with calc_orbit and backshift 2
EDGE = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                   "golden", "edge_frames.json")))
for _name, _e in EDGE.items():
    _full = not _e["calc_fast"]
    FRAMES[_name] = ("M2", dict(x=_e["x"], y=_e["y"], dx=_e["dx"], precision=_e["precision"],
                                xy_ratio=_e["xy_ratio"]),
                     dict(max_iter=_e["max_iter"], interior_detect=_e["interior_detect"],
                          calc_dzndc=_full, calc_orbit=_full, backshift=2 if _full else 0))

# z^d + c: seeded boundary-descent centres (fractalshades_amd.configs)
MN_TEST = {d: configs.Workload(f"T_MN{d}", "MN", "1e-14", 64, 64, 5000, 40, degree=d,
                               period_range=(10, 3000), seed=d) for d in (2, 3, 4, 5, 6, 7, 8)}
_mn_centers = {}


def mn_frame(d, nx=48, xr=False, **calc_over):
    if d not in _mn_centers:
        _mn_centers[d] = configs.seeded_center(MN_TEST[d])
    x, y, _, _ = _mn_centers[d]
    f = fsa.Perturbation_mandelbrot_N(None, d)
    f.zoom(precision=40, x=x, y=y, dx="1e-14", nx=nx, xy_ratio=1.0, theta_deg=0.)
    kw = dict(calc_name="t", subset=None, max_iter=5000, M_divergence=1.e3,
              epsilon_stationnary=1.e-3, BLA_eps=1.e-6, interior_detect=True)
    kw.update(calc_over)
    f.calc_std_div(**kw)
    return f


def make(name, nx=48, model_kw=None, no_newton=False, **calc_over):
    """The frame `name` at nx pixels per row; no_newton: reference orbit from
    the image centre, not a cycle (the reference's settings.no_newton)."""
    model, zoom_kw, calc_kw = FRAMES[name]
    if model == "BS" and model_kw is None:
        model_kw = {"flavor": FRAMES[name + "_flavor"]}
    if model == "M2":
        f = fsa.Perturbation_mandelbrot(None)
    elif model == "MN":
        f = fsa.Perturbation_mandelbrot_N(None, **(model_kw or {}))
    else:
        f = fsa.Perturbation_burning_ship(None, **(model_kw or {}))
    f.zoom(**dict(dict(nx=nx, xy_ratio=1.0, theta_deg=0.), **zoom_kw))
    f.no_newton = no_newton
    kw = dict(calc_name="t", subset=None, M_divergence=1.e3, epsilon_stationnary=1.e-3,
              BLA_eps=1.e-6)
    if model == "BS":
        kw.pop("epsilon_stationnary")
    kw.update(calc_kw)
    kw.update(calc_over)
    f.calc_std_div(**kw)
    return f


def frame_inputs(f, calc_name="t"):
    args = f._calc_data[calc_name]["cycle_indep_args"]
    c_pix = np.ravel(f.chunk_pixel_pos((0, f.nx, 0, f.ny)))
    return args, np.ascontiguousarray(c_pix)


def expmap_frame(name="M2_E20", nx=60, hmin=0., hmax=10., dx=None, orientation="horizontal",
                 precision=None, no_newton=False, x=None, y=None, **calc_over):
    """A test frame's location seen through the exponential map
    exp(h + i t), h in [hmin, hmax] (fractalshades_amd.projection.Expmap)."""
    from fractalshades_amd.projection import Expmap
    model, zoom_kw, calc_kw = FRAMES[name]
    zoom_kw = dict(zoom_kw, projection=Expmap(hmin, hmax, orientation=orientation))
    if dx is not None:
        zoom_kw["dx"] = dx
    if precision is not None:
        zoom_kw["precision"] = precision
    if x is not None:
        zoom_kw["x"], zoom_kw["y"] = x, y
    if model == "M2":
        f = fsa.Perturbation_mandelbrot(None)
    else:
        f = fsa.Perturbation_burning_ship(None, flavor=FRAMES[name + "_flavor"])
    f.zoom(nx=nx, xy_ratio=1.0, theta_deg=0., **zoom_kw)
    f.no_newton = no_newton
    kw = dict(calc_name="t", subset=None, M_divergence=1.e3, epsilon_stationnary=1.e-3,
              BLA_eps=1.e-6)
    if model == "BS":
        kw.pop("epsilon_stationnary")
    kw.update(calc_kw)
    kw.update(calc_over)
    f.calc_std_div(**kw)
    return f
