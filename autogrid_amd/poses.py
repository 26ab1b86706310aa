"""Ligand poses and their trilinear splat onto a map set (no torch).

Docking reads AutoGrid maps by trilinear interpolation: a ligand atom of
type t, charge q at position x scores map_t(x) + q e(x) + |q| d(x), each map
interpolated between the 8 grid points around x (x fastest, main.cpp:322).
The scores of many poses are S @ maps.ravel() with S sparse; S^T gives the
cotangent dL/dmaps of any loss of the scores.
"""
from __future__ import annotations

import numpy as np


def make_poses(st, map_of_type: dict, n_poses: int = 256, seed: int = 1005,
               atoms_range=(20, 60), sigma_q: float = 0.15, spread: float = 3.0):
    """Seeded ligand poses inside the box: per pose a centre drawn uniformly
    in the box (a margin of ``spread`` + 1 grid step from the faces) and
    atoms scattered around it (Gaussian, ``spread``/2 A), types drawn from
    ``map_of_type`` keys, charges N(0, sigma_q).  Returns a list of
    (map index [n], charge [n], position [n, 3] grid frame)."""
    rng = np.random.default_rng(seed)
    names = list(map_of_type)
    lo = np.array([-st.ne[d] * st.spacing for d in range(3)])
    hi = np.array([(st.n1[d] - 1 - st.ne[d]) * st.spacing for d in range(3)])
    margin = spread + st.spacing
    poses = []
    for _ in range(n_poses):
        n = int(rng.integers(atoms_range[0], atoms_range[1] + 1))
        c = rng.uniform(lo + margin, hi - margin)
        x = c + rng.normal(0.0, spread / 2.0, size=(n, 3))
        x = np.clip(x, lo + 1e-3, hi - st.spacing - 1e-3)
        t = rng.integers(0, len(names), size=n)
        poses.append((np.array([map_of_type[names[k]] for k in t]), rng.normal(0.0, sigma_q, size=n), x))
    return poses


def splat(st, poses):
    """(rows, cols, vals) of S: row = pose, col = map * P + point.  A map
    index of -1 (an atom without an affinity map, or no e / d map) adds no
    entries.  Atoms must lie inside the grid."""
    P = st.points
    n1x, n1y = st.n1[0], st.n1[1]
    rows, cols, vals = [], [], []
    e, d = st.elec_index, st.dsolv_index
    for k, (mi, q, x) in enumerate(poses):
        g = (x + np.array([st.ne[0], st.ne[1], st.ne[2]]) * st.spacing) / st.spacing
        i0 = np.floor(g).astype(np.int64)
        f = g - i0
        for dz in (0, 1):
            for dy in (0, 1):
                for dx in (0, 1):
                    w = (f[:, 0] if dx else 1 - f[:, 0]) * (f[:, 1] if dy else 1 - f[:, 1]) * \
                        (f[:, 2] if dz else 1 - f[:, 2])
                    pt = (i0[:, 0] + dx) + n1x * ((i0[:, 1] + dy) + n1y * (i0[:, 2] + dz))
                    for m, scale in ((mi, 1.0), (e, q), (d, np.abs(q))):
                        m = np.broadcast_to(np.asarray(m, np.int64), pt.shape)
                        keep = m >= 0          # -1: no affinity map / no e or d map
                        rows.append(np.full(int(keep.sum()), k))
                        cols.append(m[keep] * P + pt[keep])
                        vals.append(np.broadcast_to(w * scale, pt.shape)[keep])
    return np.concatenate(rows), np.concatenate(cols), np.concatenate(vals)


def cotangent(S, n_cols: int, dscore) -> np.ndarray:
    """S^T dscore: dL/dmaps (flat) for dL/dscore."""
    r, c, v = S
    out = np.zeros(n_cols)
    np.add.at(out, c, v * np.asarray(dscore)[r])
    return out
