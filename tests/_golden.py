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
