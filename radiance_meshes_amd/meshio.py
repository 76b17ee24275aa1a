"""On-disk mesh format for loading device-resident meshes (serving, evaluation).

The reference writes a trained mesh with `BaseModel.save2ply` (models/base_model.py:107-151):
a binary PLY with element `vertex` (x, y, z) and element `tetrahedron` (indices, s, grd_x/y/z,
sh_0_r/g/b = RGB2SH of the raw base colour at vertex 0, sh_1.._K_r/g/b the higher spherical-
harmonic coefficients). Its writer (tinyplypy) is not available here and its reader
(`Model.load_ply`, called by evaluate.py:30) is not defined in the reference, so the tetrahedron
indices are written here as the usual PLY list property (`property list uchar int indices`);
`read_ply` also accepts them as four scalar int properties.

`view_features` turns a loaded mesh into the per-tet feature rows the compositor takes for one
camera, as the reference's `activate_output` does (utils/model_util.py:93-105): colour =
softplus_10(0.5 + SH(direction from the camera to the tet's centroid)), features
[density, colour, gradient].
"""
from __future__ import annotations

import math
import os

import numpy as np

_PLY_TYPES = {
    "char": "i1", "int8": "i1", "uchar": "u1", "uint8": "u1",
    "short": "i2", "int16": "i2", "ushort": "u2", "uint16": "u2",
    "int": "i4", "int32": "i4", "uint": "u4", "uint32": "u4",
    "float": "f4", "float32": "f4", "double": "f8", "float64": "f8",
}
_NP_TO_PLY = {"i1": "char", "u1": "uchar", "i2": "short", "u2": "ushort", "i4": "int", "u4": "uint",
              "f4": "float", "f8": "double"}

SH_C0 = 0.28209479177387814


def write_ply(path, elements: dict):
    """Binary little-endian PLY. elements: {name: {prop: array}}; a 2-D integer array becomes a
    list property (uchar count, int items), 1-D arrays scalar properties of their dtype."""
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    header = ["ply", "format binary_little_endian 1.0"]
    blocks = []
    for name, props in elements.items():
        n = None
        cols = []
        for pname, a in props.items():
            a = np.asarray(a)
            m = a.shape[0]
            if n is None:
                n = m
            if m != n:
                raise ValueError(f"element {name}: property {pname} has {m} rows, expected {n}")
            cols.append((pname, a))
        header.append(f"element {name} {n}")
        fields = []
        for pname, a in cols:
            if a.ndim == 2:
                if a.shape[1] > 255:
                    raise ValueError("list properties hold at most 255 items")
                header.append(f"property list uchar int {pname}")
                fields += [(pname + "__n", "u1"), (pname, "<i4", (a.shape[1],))]
            else:
                t = a.dtype.str.lstrip("<>|=")
                if t not in _NP_TO_PLY:
                    raise ValueError(f"unsupported dtype {a.dtype} for {pname}")
                header.append(f"property {_NP_TO_PLY[t]} {pname}")
                fields.append((pname, "<" + t))
        rec = np.zeros(n, dtype=np.dtype(fields))
        for pname, a in cols:
            if a.ndim == 2:
                rec[pname + "__n"] = a.shape[1]
                rec[pname] = a
            else:
                rec[pname] = a
        blocks.append(rec.tobytes())
    header.append("end_header")
    with open(path, "wb") as f:
        f.write(("\n".join(header) + "\n").encode("ascii"))
        for b in blocks:
            f.write(b)


def read_ply(path) -> dict:
    """{element: {property: array}} of an ASCII or binary (little/big endian) PLY. List
    properties of a constant length come back as 2-D arrays, others as lists of arrays."""
    with open(path, "rb") as f:
        data = f.read()
    end = data.find(b"end_header")
    if not data.startswith(b"ply") or end < 0:
        raise ValueError(f"{path}: not a PLY file")
    nl = data.find(b"\n", end)
    header = data[:end].decode("ascii").splitlines()
    body = data[nl + 1:]
    fmt = None
    elements = []   # (name, count, [(prop, kind, type, count_type)])
    for line in header[1:]:
        w = line.split()
        if not w or w[0] in ("comment", "obj_info"):
            continue
        if w[0] == "format":
            fmt = w[1]
        elif w[0] == "element":
            elements.append((w[1], int(w[2]), []))
        elif w[0] == "property":
            if w[1] == "list":
                elements[-1][2].append((w[4], "list", _PLY_TYPES[w[3]], _PLY_TYPES[w[2]]))
            else:
                elements[-1][2].append((w[2], "scalar", _PLY_TYPES[w[1]], None))
    if fmt not in ("ascii", "binary_little_endian", "binary_big_endian"):
        raise ValueError(f"{path}: unsupported PLY format {fmt}")
    out = {}
    if fmt == "ascii":
        tokens = body.split()
        pos = 0
        for name, n, props in elements:
            cols = {p[0]: [] for p in props}
            for _ in range(n):
                for pname, kind, t, ct in props:
                    if kind == "list":
                        k = int(tokens[pos]); pos += 1
                        cols[pname].append(np.array(tokens[pos:pos + k], dtype=t)); pos += k
                    else:
                        cols[pname].append(np.array(tokens[pos], dtype=t)); pos += 1
            out[name] = {p: _stack(v, kind == "list") for (p, kind, _, _), v in zip(props, cols.values())}
        return out
    bo = "<" if fmt == "binary_little_endian" else ">"
    pos = 0
    for name, n, props in elements:
        if n == 0:
            out[name] = {p[0]: np.zeros(0) for p in props}
            continue
        # lists of constant length read as one structured array, found from the first row
        fields, probe = [], pos
        for pname, kind, t, ct in props:
            if kind == "list":
                cnt = int(np.frombuffer(body, dtype=bo + ct, count=1, offset=probe)[0])
                fields += [(pname + "__n", bo + ct), (pname, bo + t, (cnt,))]
                probe += np.dtype(ct).itemsize + cnt * np.dtype(t).itemsize
            else:
                fields.append((pname, bo + t))
                probe += np.dtype(t).itemsize
        dt = np.dtype(fields)
        if pos + n * dt.itemsize <= len(body):
            rec = np.frombuffer(body, dtype=dt, count=n, offset=pos)
            ok = all(np.all(rec[p + "__n"] == rec.dtype[p].shape[0]) for p, kind, _, _ in props if kind == "list")
        else:
            ok = False
        if ok:
            out[name] = {p: np.array(rec[p]) for p, _, _, _ in props}
            pos += n * dt.itemsize
            continue
        cols = {p[0]: [] for p in props}   # variable-length lists: row by row
        for _ in range(n):
            for pname, kind, t, ct in props:
                if kind == "list":
                    k = int(np.frombuffer(body, dtype=bo + ct, count=1, offset=pos)[0])
                    pos += np.dtype(ct).itemsize
                    cols[pname].append(np.frombuffer(body, dtype=bo + t, count=k, offset=pos).copy())
                    pos += k * np.dtype(t).itemsize
                else:
                    cols[pname].append(np.frombuffer(body, dtype=bo + t, count=1, offset=pos)[0])
                    pos += np.dtype(t).itemsize
        out[name] = {p: _stack(v, kind == "list") for (p, kind, _, _), v in zip(props, cols.values())}
    return out


def _stack(values, is_list):
    if not is_list:
        return np.array(values)
    lens = {len(v) for v in values}
    return np.stack(values) if len(lens) == 1 else values


def rgb_to_sh(rgb):
    return (np.asarray(rgb) - 0.5) / SH_C0


def save_mesh_ply(path, vertices, indices, density, rgb, grd, sh_rest=None):
    """The reference's save2ply layout (models/base_model.py:107-151): vertex xyz; tetrahedron
    indices, s (density), grd_x/y/z (the normalised colour gradient), sh_0_{r,g,b} = RGB2SH(rgb)
    with rgb the field's raw colour at the tet's anchor point (what compute_features(offset=True)
    returns, models/ingp_color.py:487-512: the centroid when the model was trained with
    ablate_circumsphere, the circumcenter otherwise), and sh_{i}_{r,g,b}, i = 1..K, from
    sh_rest [T, K, 3]."""
    v = np.asarray(vertices, np.float32)
    tet = {"indices": np.asarray(indices, np.int32), "s": np.asarray(density, np.float32).reshape(-1)}
    g = np.asarray(grd, np.float32).reshape(-1, 3)
    for i, co in enumerate("xyz"):
        tet[f"grd_{co}"] = np.ascontiguousarray(g[:, i])
    sh0 = rgb_to_sh(np.asarray(rgb, np.float32).reshape(-1, 3)).astype(np.float32)
    for i, ch in enumerate("rgb"):
        tet[f"sh_0_{ch}"] = np.ascontiguousarray(sh0[:, i])
    if sh_rest is not None:
        sr = np.asarray(sh_rest, np.float32)
        for k in range(sr.shape[1]):
            for i, ch in enumerate("rgb"):
                tet[f"sh_{k + 1}_{ch}"] = np.ascontiguousarray(sr[:, k, i])
    write_ply(path, {"vertex": {"x": v[:, 0], "y": v[:, 1], "z": v[:, 2]}, "tetrahedron": tet})


def load_mesh_ply(path) -> dict:
    """vertices [V,3] f32, indices [T,4] i32, density [T], grd [T,3], sh0 [T,3], sh_rest [T,K,3]."""
    d = read_ply(path)
    vx = d["vertex"]
    tet = d["tetrahedron"]
    if "indices" in tet:
        idx = np.asarray(tet["indices"], np.int32)
    else:
        idx = np.stack([tet[f"vertex_index_{k}"] for k in range(4)], 1).astype(np.int32)
    if idx.ndim != 2 or idx.shape[1] != 4:
        raise ValueError(f"{path}: tetrahedron indices must have 4 entries per row")
    K = 0
    while f"sh_{K + 1}_r" in tet:
        K += 1
    sh_rest = np.stack([np.stack([tet[f"sh_{k + 1}_{c}"] for c in "rgb"], 1) for k in range(K)], 1) \
        if K else np.zeros((idx.shape[0], 0, 3), np.float32)
    return {
        "vertices": np.stack([vx["x"], vx["y"], vx["z"]], 1).astype(np.float32),
        "indices": idx,
        "density": np.asarray(tet["s"], np.float32),
        "grd": np.stack([tet[f"grd_{c}"] for c in "xyz"], 1).astype(np.float32),
        "sh0": np.stack([tet[f"sh_0_{c}"] for c in "rgb"], 1).astype(np.float32),
        "sh_rest": sh_rest.astype(np.float32),
    }


def sh_basis(dirs, deg: int):
    """Real spherical-harmonic basis at unit directions [N,3] for degrees 1..deg (deg <= 3):
    [N, (deg+1)^2 - 1], ordered by degree and m = -l..l, without the constant term. Each function
    is sqrt(2) Im (m < 0), Re (m = 0) or sqrt(2) Re (m > 0) of the complex harmonic Y_l^|m| with the
    Condon-Shortley phase: the sign convention of the graphics SH tables (utils/eval_sh_py.py:137-195),
    Y_1,-1 = -sqrt(3 / 4 pi) y. Coefficients are the closed-form normalisations."""
    import torch
    x, y, z = dirs[:, 0], dirs[:, 1], dirs[:, 2]
    cols = []
    if deg >= 1:
        c1 = math.sqrt(3.0 / (4 * math.pi))
        cols += [-c1 * y, c1 * z, -c1 * x]
    if deg >= 2:
        xx, yy, zz = x * x, y * y, z * z
        c2a = 0.5 * math.sqrt(15.0 / math.pi)
        c2b = 0.25 * math.sqrt(5.0 / math.pi)
        c2c = 0.25 * math.sqrt(15.0 / math.pi)
        cols += [c2a * x * y, -c2a * y * z, c2b * (2 * zz - xx - yy), -c2a * x * z, c2c * (xx - yy)]
    if deg >= 3:
        c3a = 0.25 * math.sqrt(35.0 / (2 * math.pi))
        c3b = 0.5 * math.sqrt(105.0 / math.pi)
        c3c = 0.25 * math.sqrt(21.0 / (2 * math.pi))
        c3d = 0.25 * math.sqrt(7.0 / math.pi)
        c3e = 0.25 * math.sqrt(105.0 / math.pi)
        cols += [-c3a * y * (3 * xx - yy), c3b * x * y * z, -c3c * y * (4 * zz - xx - yy),
                 c3d * z * (2 * zz - 3 * xx - 3 * yy), -c3c * x * (4 * zz - xx - yy),
                 c3e * z * (xx - yy), -c3a * x * (xx - 3 * yy)]
    if deg > 3:
        raise ValueError("SH degree above 3 is not supported")
    return torch.stack(cols, 1) if cols else torch.zeros((dirs.shape[0], 0), dtype=dirs.dtype, device=dirs.device)


def tet_anchor(verts, idx, anchor: str):
    """The point a tet's colour field is anchored at: "centroid" (the mean of its vertices) or
    "circumcenter" (the textbook circumsphere centre, in FP64 as calculate_circumcenters_torch
    evaluates it on double vertices, models/ingp_color.py:320)."""
    import torch
    tets = verts[idx]
    if anchor == "centroid":
        return tets.mean(dim=1)
    if anchor != "circumcenter":
        raise ValueError(f"anchor must be 'centroid' or 'circumcenter', got {anchor!r}")
    t = tets.double()
    a, b, c = t[:, 1] - t[:, 0], t[:, 2] - t[:, 0], t[:, 3] - t[:, 0]
    num = ((a * a).sum(1, keepdim=True) * torch.cross(b, c, dim=1) + (b * b).sum(1, keepdim=True)
           * torch.cross(c, a, dim=1) + (c * c).sum(1, keepdim=True) * torch.cross(a, b, dim=1))
    den = 2 * (a * torch.cross(b, c, dim=1)).sum(1, keepdim=True)
    return (t[:, 0] + num / den).to(verts.dtype)


def view_features(mesh, cam_pos, sh_deg: int = 3, device="cuda", anchor: str = "centroid"):
    """Per-tet feature rows [T,7] = [density, colour, gradient] for a camera at cam_pos, from a
    mesh dict of load_mesh_ply (tensors or arrays), as activate_output (utils/model_util.py:93-105)
    forms them: the colour field is SH2RGB(sh0) + grd . (p - anchor) (offset_normalize,
    :81-91), evaluated at vertex 0 and view-dependent through the higher SH bands:
    colour = softplus_10(SH2RGB(sh0) + grd . (v0 - anchor) + sum_k Y_k(dir) sh_k), dir the unit
    vector from the camera to the tet's centroid. anchor: "centroid" -- where the reference's
    model anchors the field when trained with ablate_circumsphere (train.py:145, its default) --
    or "circumcenter"."""
    import torch
    t = {k: torch.as_tensor(v, device=device) for k, v in mesh.items()}
    verts, idx = t["vertices"].float(), t["indices"].long()
    cen = verts[idx].mean(dim=1)
    cp = torch.as_tensor(cam_pos, device=device, dtype=torch.float32).reshape(1, 3)
    d = torch.nn.functional.normalize(cen - cp, dim=1, eps=1e-8)
    off = verts[idx[:, 0]] - tet_anchor(verts, idx, anchor)
    col = SH_C0 * t["sh0"].float() + 0.5 + (t["grd"].float() * off).sum(1, keepdim=True)
    K = t["sh_rest"].shape[1]
    deg = min(sh_deg, int(round(math.sqrt(K + 1))) - 1)
    if deg > 0:
        B = sh_basis(d, deg)                      # [T, (deg+1)^2 - 1]
        col = col + torch.einsum("tk,tkc->tc", B, t["sh_rest"][:, :B.shape[1]].float())
    col = torch.nn.functional.softplus(col, beta=10)
    return torch.cat([t["density"].float().reshape(-1, 1), col, t["grd"].float()], 1).contiguous()
