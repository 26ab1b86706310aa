"""Seeded synthetic receptors and GPFs for the grid-map workloads.

A receptor is a self-avoiding chain of heavy atoms (bonds 1.2-1.55 A, atoms
two apart at least 2.0 A, atoms three or more apart at least 2.9 A) packed in
a sphere of the given radius.  Heavy types are drawn C 60 : N 15 : OA 18 :
SA 1; 20% of N/OA atoms carry a donor hydrogen HD placed 1.0 A from the
parent, directly after it in atom order and with no other atom within 1.5 A.
Charges are Gaussian (sigma 0.2 e) shifted to a zero total.  The PDBQT uses
the fixed columns autogrid4 reads (main.cpp:881-902): x/y/z in columns 31-54
(%8.3f), the charge in 71-76 (%6.3f) and the type in 78-79.

Workloads (grid centred on the receptor centroid, spacing 0.375 A,
AD4.1_bound parameters, smooth 0.5, distance-dependent dielectric):

* ``small``  -- 1,000 atoms in a 28 A sphere, npts 40 (41^3 points), maps
  A C HD N NA OA + e + d.
* ``full``   -- 5,000 atoms in a 45 A sphere, npts 126 (127^3 points), maps
  A C N NA OA S SA HD F Cl Br I P + e + d.
* ``large``  -- 20,000 atoms in an 80 A sphere, npts 256 (257^3 points),
  same maps as ``full``.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass

import numpy as np

SMALL_TYPES = ["A", "C", "HD", "N", "NA", "OA"]
FULL_TYPES = ["A", "C", "N", "NA", "OA", "S", "SA", "HD", "F", "Cl", "Br", "I", "P"]


@dataclass
class Workload:
    name: str
    n_atoms: int
    radius: float
    npts: int
    ligand_types: list
    seed: int


WORKLOADS = {
    "small": Workload("small", 1000, 28.0, 40, SMALL_TYPES, 1002),
    "full": Workload("full", 5000, 45.0, 126, FULL_TYPES, 1003),
    "large": Workload("large", 20000, 80.0, 256, FULL_TYPES, 1004),
}


class _Hash:
    """Uniform spatial hash for the placement checks."""

    def __init__(self, cell: float):
        self.cell = cell
        self.d: dict = {}

    def key(self, p):
        return (int(math.floor(p[0] / self.cell)), int(math.floor(p[1] / self.cell)), int(math.floor(p[2] / self.cell)))

    def add(self, i, p):
        self.d.setdefault(self.key(p), []).append(i)

    def near(self, p):
        kx, ky, kz = self.key(p)
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    yield from self.d.get((kx + dx, ky + dy, kz + dz), ())


def make_receptor(n_atoms: int, radius: float, seed: int, center=(0.0, 0.0, 0.0)):
    """(coords [n,3], types [n] str, charges [n]) of a seeded receptor."""
    rng = np.random.default_rng(seed)
    heavy_types = np.array(["C", "N", "OA", "SA"])
    heavy_w = np.array([60.0, 15.0, 18.0, 1.0])
    heavy_w /= heavy_w.sum()
    pos: list = []
    typ: list = []
    heavy_index: list = []     # index into pos of each heavy atom, chain order
    hh = _Hash(3.0)            # heavy atoms
    ha = _Hash(3.0)            # all atoms
    r2 = radius * radius

    def free_heavy(p, chain_pos):
        for j in hh.near(p):
            d = pos[j] - p
            dd = float(d @ d)
            cj = heavy_chain_pos[j]
            gap = chain_pos - cj
            if gap >= 3 and dd < 2.9 * 2.9:
                return False
            if gap == 2 and dd < 2.0 * 2.0:
                return False
            if dd < 1.0:
                return False
        for j in ha.near(p):
            if typ[j] == "HD":
                d = pos[j] - p
                if float(d @ d) < 1.5 * 1.5:
                    return False
        return True

    heavy_chain_pos: dict = {}
    chain_pos = 0
    prev = None
    while len(pos) < n_atoms:
        placed = None
        if prev is not None:
            for _ in range(60):
                v = rng.normal(size=3)
                v /= np.linalg.norm(v)
                p = prev + v * rng.uniform(1.2, 1.55)
                if float(p @ p) <= r2 and free_heavy(p, chain_pos):
                    placed = p
                    break
        if placed is None:
            # start a new chain segment at a random free point
            chain_pos += 10
            for _ in range(100000):
                p = rng.uniform(-radius, radius, size=3)
                if float(p @ p) <= r2 and free_heavy(p, chain_pos):
                    placed = p
                    break
            if placed is None:
                raise RuntimeError("receptor generator could not place an atom")
        t = str(rng.choice(heavy_types, p=heavy_w))
        i = len(pos)
        pos.append(placed)
        typ.append(t)
        heavy_chain_pos[i] = chain_pos
        hh.add(i, placed)
        ha.add(i, placed)
        heavy_index.append(i)
        prev = placed
        chain_pos += 1
        if t in ("N", "OA") and len(pos) < n_atoms and rng.uniform() < 0.2:
            for _ in range(60):
                v = rng.normal(size=3)
                v /= np.linalg.norm(v)
                h = placed + v * 1.0
                ok = True
                for j in ha.near(h):
                    if j == i:
                        continue
                    d = pos[j] - h
                    if float(d @ d) < 1.5 * 1.5:
                        ok = False
                        break
                if ok:
                    k = len(pos)
                    pos.append(h)
                    typ.append("HD")
                    ha.add(k, h)
                    break
    coords = np.array(pos) + np.asarray(center, float)
    q = rng.normal(0.0, 0.2, size=n_atoms)
    q -= q.mean()
    return coords, np.array(typ), q


def write_pdbqt(path: str, coords, types, charges) -> None:
    with open(path, "w") as f:
        for i, (p, t, q) in enumerate(zip(coords, types, charges)):
            name = {"C": "C", "N": "N", "OA": "O", "SA": "S", "HD": "H"}.get(t, t[:1])
            f.write("ATOM  %5d %-4s %3s A%4d    %8.3f%8.3f%8.3f  1.00  0.00    %6.3f %-2s\n"
                    % (i % 100000, name, "UNK", (i // 10) % 10000, p[0], p[1], p[2], q, t))


def read_pdbqt_coords(path: str):
    xyz = []
    for line in open(path):
        if line.startswith(("ATOM  ", "HETATM")):
            xyz.append((float(line[30:38]), float(line[38:46]), float(line[46:54])))
    return np.array(xyz)


def write_gpf(path: str, stem: str, receptor: str, rec_types, lig_types, npts, center,
              spacing=0.375, extra=()) -> list:
    """Write a GPF; returns the map file names in slot order."""
    if isinstance(npts, int):
        npts = (npts, npts, npts)
    maps = [f"{stem}.{t}.map" for t in lig_types] + [f"{stem}.e.map", f"{stem}.d.map"]
    lines = [f"npts {npts[0]} {npts[1]} {npts[2]}",
             f"gridfld {stem}.maps.fld",
             f"spacing {spacing}",
             "receptor_types " + " ".join(rec_types),
             "ligand_types " + " ".join(lig_types),
             f"receptor {receptor}",
             "gridcenter %.3f %.3f %.3f" % tuple(center),
             "smooth 0.5"]
    lines += list(extra)
    lines += [f"map {m}" for m in maps[:-2]]
    lines += [f"elecmap {maps[-2]}", f"dsolvmap {maps[-1]}", "dielectric -0.1465"]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return maps


def make_workload(name: str, outdir: str, npts=None, z_npts=None, center_offset=(11.5, -7.25, 3.125)) -> dict:
    """Generate receptor + GPF for a named workload into outdir.

    npts overrides the cube size; z_npts (if set) shrinks only the z axis so a
    slab of the workload can be timed on its own."""
    w = WORKLOADS[name]
    os.makedirs(outdir, exist_ok=True)
    coords, types, q = make_receptor(w.n_atoms, w.radius, w.seed, center=center_offset)
    rec = f"{name}.pdbqt"
    write_pdbqt(os.path.join(outdir, rec), coords, types, q)
    written = read_pdbqt_coords(os.path.join(outdir, rec))
    centroid = written.mean(axis=0)
    present = []
    for t in types:
        if t not in present:
            present.append(t)
    n = w.npts if npts is None else npts
    nxyz = (n, n, n if z_npts is None else z_npts)
    gpf = f"{name}.gpf"
    maps = write_gpf(os.path.join(outdir, gpf), name, rec, present, w.ligand_types, nxyz, centroid)
    return {"dir": outdir, "gpf": gpf, "receptor": rec, "maps": maps, "n_atoms": w.n_atoms,
            "npts": nxyz, "center": centroid, "types": present}
