"""GPU parity of the fused per-tet feature activation (rm_activate_fwd/bwd) against the FP64
oracle: features, and gradients against central differences of the oracle with the vertex-0
colour offset held fixed (the reference detaches it)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from tests.test_activate_oracle import _inputs


@pytest.mark.parametrize("deg", [0, 1, 3])
def test_activate_forward_and_backward(oracle_lib, deg):
    import torch
    from radiance_meshes_amd.activate import activate_output
    O = oracle_lib
    a = _inputs(T=300, V=400, seed=deg)
    print(f"activate deg={deg}", flush=True)
    t = {k: torch.tensor(np.asarray(v, np.float32) if k != "indices" else np.asarray(v, np.int64), device="cuda")
         for k, v in a.items()}
    for k in ("density", "rgb", "grd", "sh", "circumcenters", "vertices"):
        t[k].requires_grad_(True)
    f = activate_output(t["camera_center"], t["density"], t["rgb"], t["grd"], t["sh"], t["indices"],
                        t["circumcenters"], t["vertices"], deg, 3)
    ref = O.activate_output(deg=deg, **a)
    err = np.abs(f.detach().cpu().numpy() - ref) / np.maximum(1, np.abs(ref))
    print(f"features: max rel err vs FP64 {err.max():.2e}")
    assert err.max() <= 2e-6
    rng = np.random.default_rng(7)
    G = rng.standard_normal(ref.shape)
    (f * torch.from_numpy(G).float().cuda()).sum().backward()
    torch.cuda.synchronize()
    h = 1e-6
    worst = 0.0
    off0 = np.asarray(a["vertices"], np.float64)[a["indices"][:, 0]] - a["circumcenters"]   # held fixed (detached)
    for name, nrow in (("density", 20), ("rgb", 30), ("grd", 30), ("sh", 40 if deg else 0),
                       ("circumcenters", 30), ("vertices", 40)):
        g = t[name].grad
        if nrow == 0:
            continue
        assert g is not None, name
        g = g.cpu().numpy().astype(np.float64)
        arr = np.asarray(a[name], np.float64)
        flat = arr.reshape(-1)
        rows = []
        for i in rng.choice(flat.size, nrow, replace=False):
            p, m = flat.copy(), flat.copy()
            p[i] += h
            m[i] -= h
            ap, am = dict(a), dict(a)
            ap[name], am[name] = p.reshape(arr.shape), m.reshape(arr.shape)
            fd = ((O.activate_output(deg=deg, offset=off0, **ap) - O.activate_output(deg=deg, offset=off0, **am))
                  * G).sum() / (2 * h)
            rows.append((g.reshape(-1)[i], fd))
        r = np.array(rows)
        scale = np.abs(r[:, 1]).max()
        e = np.abs(r[:, 0] - r[:, 1]) / np.maximum(np.abs(r[:, 1]), 1e-2 * scale + 1e-12)
        worst = max(worst, e.max())
        print(f"{name}: max rel err vs FP64 central differences {e.max():.2e}")
        assert e.max() <= 1e-3, name
