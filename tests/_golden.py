"""Loads a golden fixture (tests/golden/*.npz) as inputs + expected oracle outputs."""
import os

import numpy as np

from radiance_meshes_amd import scenes

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    z = dict(np.load(os.path.join(HERE, f"{name}.npz")))
    H, W = (int(v) for v in z["hw"])
    z["camera"] = scenes.Camera(z["w2v"], z["K"], z["cam_pos"], H, W, 1.0, 1.0)
    z["tile"] = int(z["tile"])
    z["min_t"] = float(z["min_t"])
    return z


FAMILIES = ("face_view", "face_view_in", "frustum", "multi_views")


def _finish(z):
    H, W = (int(v) for v in z["hw"])
    z["camera"] = scenes.Camera(z["w2v"], z["K"], z["cam_pos"], H, W, 1.0, 1.0)
    z["tile"] = int(z["tile"])
    z["min_t"] = float(z["min_t"])
    return z


def load_family(name):
    """[(case, fixture dict)] of a multi-case fixture file (keys '<case>/<field>')."""
    raw = np.load(os.path.join(HERE, f"{name}.npz"))
    cases = {}
    for key in raw.files:
        c, f = key.split("/", 1)
        cases.setdefault(c, {})[f] = raw[key]
    return [(c, _finish(d)) for c, d in cases.items()]


def family_cases():
    """(family, case) ids of every multi-case fixture, for parametrisation."""
    out = []
    for fam in FAMILIES:
        raw = np.load(os.path.join(HERE, f"{fam}.npz"))
        out += sorted({(fam, k.split("/", 1)[0]) for k in raw.files})
    return out


def load_case(family, case):
    return dict(load_family(family))[case]
