"""Gaussian PLY files (the format of GaussianIO.save_as_ply, GaussianIO.py:9-45).

One `vertex` element, every property float32, in this order:
``x y z``, ``sh_coefficients_dc_0..2`` (r, g, b), ``sh_coefficients_ac_i``
for i = channel * K + coefficient (K = (degree + 1)^2 - 1: the reference
flattens its [N, K, 3] tensor channel-major, misc_utils.py:41-43),
``opacity`` (logit), ``svec_0..2`` (log-scales), ``quaternion_0..3``
(w, x, y, z).  Files are binary little-endian, as plyfile writes them on
x86; big-endian and ASCII files are read too.

Reading inverts the writer exactly.  The reference's reader does not: it
unflattens the AC block coefficient-major (GaussianIO.py:59-63, i % 3 as the
channel) and is called with an argument it does not take
(gaussians_handler.py:92); this module follows the file as written.
"""
from __future__ import annotations

import os

import numpy as np

_FLOAT_NAMES = {"float", "float32"}


def _props(K: int) -> list:
    names = ["x", "y", "z"] + [f"sh_coefficients_dc_{i}" for i in range(3)]
    names += [f"sh_coefficients_ac_{i}" for i in range(3 * K)]
    names += ["opacity"] + [f"svec_{i}" for i in range(3)] + [f"quaternion_{i}" for i in range(4)]
    return names


def _np(t) -> np.ndarray:
    if hasattr(t, "detach"):
        t = t.detach().cpu().numpy()
    return np.asarray(t, dtype=np.float32)


def save_ply(path: str, means, sh_dc, sh_ac, opacity, svec, quaternion) -> None:
    """means [N,3], sh_dc [N,1,3], sh_ac [N,K,3], opacity [N] or [N,1] (logits),
    svec [N,3] (log-scales), quaternion [N,4]; numpy arrays or torch tensors."""
    means, sh_dc, sh_ac = _np(means), _np(sh_dc), _np(sh_ac)
    opacity, svec, quaternion = _np(opacity).reshape(-1), _np(svec), _np(quaternion)
    n = means.shape[0]
    K = sh_ac.shape[1] if sh_ac.size else 0
    if sh_dc.shape != (n, 1, 3) or sh_ac.shape not in ((n, K, 3),) or opacity.shape != (n,) or \
            svec.shape != (n, 3) or quaternion.shape != (n, 4) or means.shape != (n, 3):
        raise ValueError("save_ply: inconsistent shapes")
    names = _props(K)
    rec = np.empty(n, dtype=np.dtype([(nm, "<f4") for nm in names]))
    for i, c in enumerate("xyz"):
        rec[c] = means[:, i]
    for i in range(3):
        rec[f"sh_coefficients_dc_{i}"] = sh_dc[:, 0, i]
    ac_flat = sh_ac.transpose(0, 2, 1).reshape(n, 3 * K)  # channel-major
    for i in range(3 * K):
        rec[f"sh_coefficients_ac_{i}"] = ac_flat[:, i]
    rec["opacity"] = opacity
    for i in range(3):
        rec[f"svec_{i}"] = svec[:, i]
    for i in range(4):
        rec[f"quaternion_{i}"] = quaternion[:, i]
    header = ["ply", "format binary_little_endian 1.0", f"element vertex {n}"]
    header += [f"property float {nm}" for nm in names]
    header += ["end_header"]
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    with open(path, "wb") as f:
        f.write(("\n".join(header) + "\n").encode("ascii"))
        f.write(rec.tobytes())


def _read_header(f):
    lines = []
    while True:
        line = f.readline()
        if not line:
            raise ValueError("PLY header not terminated")
        line = line.decode("ascii").strip()
        lines.append(line)
        if line == "end_header":
            break
    if not lines or lines[0] != "ply":
        raise ValueError("not a PLY file")
    fmt, n, props, in_vertex = None, None, [], False
    for line in lines[1:]:
        tok = line.split()
        if not tok or tok[0] in ("comment", "obj_info"):
            continue
        if tok[0] == "format":
            fmt = tok[1]
        elif tok[0] == "element":
            in_vertex = tok[1] == "vertex"
            if in_vertex:
                n = int(tok[2])
            elif n is None and int(tok[2]) > 0:
                # its records would precede the vertices in the body
                raise ValueError(f"element {tok[1]!r} before 'vertex' is not supported")
        elif tok[0] == "property" and in_vertex:
            if tok[1] == "list":
                raise ValueError("list properties are not part of the Gaussian format")
            if tok[1] not in _FLOAT_NAMES:
                raise ValueError(f"property {tok[2]}: type {tok[1]}, expected float")
            props.append(tok[2])
    if fmt not in ("binary_little_endian", "binary_big_endian", "ascii") or n is None:
        raise ValueError(f"unsupported PLY format {fmt!r}")
    return fmt, n, props


def load_ply(path: str) -> dict:
    """{'means' [N,3], 'sh_dc' [N,1,3], 'sh_ac' [N,K,3], 'opacity' [N],
    'svec' [N,3], 'quaternion' [N,4]} as float32 numpy arrays."""
    with open(path, "rb") as f:
        fmt, n, props = _read_header(f)
        if fmt != "ascii":
            e = "<f4" if fmt == "binary_little_endian" else ">f4"
            body = f.read(4 * len(props) * n)
            if len(body) != 4 * len(props) * n:
                raise ValueError("PLY body shorter than its header says")
            rec = np.frombuffer(body, dtype=np.dtype([(p, e) for p in props]), count=n)
            col = lambda p: np.asarray(rec[p], np.float32)
        else:
            data = np.loadtxt(f, dtype=np.float32, ndmin=2, max_rows=n) if n else \
                np.zeros((0, len(props)), np.float32)
            idx = {p: i for i, p in enumerate(props)}
            col = lambda p: data[:, idx[p]]
    need = ["x", "y", "z", "opacity"] + [f"sh_coefficients_dc_{i}" for i in range(3)] + \
        [f"svec_{i}" for i in range(3)] + [f"quaternion_{i}" for i in range(4)]
    missing = [p for p in need if p not in props]
    if missing:
        raise ValueError(f"PLY lacks Gaussian properties {missing}")
    nac = sum(1 for p in props if p.startswith("sh_coefficients_ac_"))
    if nac % 3:
        raise ValueError("AC coefficient count is not a multiple of 3")
    K = nac // 3
    means = np.stack([col(c) for c in "xyz"], 1)
    sh_dc = np.stack([col(f"sh_coefficients_dc_{i}") for i in range(3)], 1)[:, None, :]
    ac_flat = np.stack([col(f"sh_coefficients_ac_{i}") for i in range(3 * K)], 1) if K else \
        np.zeros((n, 0), np.float32)
    sh_ac = ac_flat.reshape(n, 3, K).transpose(0, 2, 1).copy()
    return {"means": means, "sh_dc": sh_dc, "sh_ac": sh_ac, "opacity": col("opacity").copy(),
            "svec": np.stack([col(f"svec_{i}") for i in range(3)], 1),
            "quaternion": np.stack([col(f"quaternion_{i}") for i in range(4)], 1)}
