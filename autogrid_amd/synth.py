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
* ``analytic`` -- the ``full`` receptor, npts 64 (65^3 points), maps
  A C HD N OA + e + d (the analytic-mode workload).
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
    # the full workload's receptor on a 65^3 box for the analytic mode
    "analytic": Workload("analytic", 5000, 45.0, 64, ["A", "C", "HD", "N", "OA"], 1003),
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

    npts overrides the cube size (an int, or an (x, y, z) tuple for a box);
    z_npts (if set) shrinks only the z axis so a slab of the workload can be
    timed on its own."""
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
    nxyz = tuple(n) if isinstance(n, (tuple, list)) else (n, n, n)
    if z_npts is not None:
        nxyz = (nxyz[0], nxyz[1], z_npts)
    gpf = f"{name}.gpf"
    maps = write_gpf(os.path.join(outdir, gpf), name, rec, present, w.ligand_types, nxyz, centroid)
    return {"dir": outdir, "gpf": gpf, "receptor": rec, "maps": maps, "n_atoms": w.n_atoms,
            "npts": nxyz, "center": centroid, "types": present}


# ---------------------------------------------------------------------------
# Ensemble of receptor conformations (the batched workload)
# ---------------------------------------------------------------------------

@dataclass
class Ensemble:
    """Base receptor (parsed by the engine's front end) and its conformations."""
    base: object                  # PreparedState of the base receptor (tables, map set, pairs)
    center: np.ndarray            # grid centre of the base GPF (its atoms are stored relative to it)
    confs: list                   # dicts: n_atoms, coord (absolute, [n,3]), npts, center
    lig_types: list               # map types in slot order
    rec_types: list               # receptor types in rec-index order


def make_ensemble(outdir: str, n_conf: int = 64, seed: int = 1005, base_atoms: int = 12000,
                  n_range=(1500, 12000), npts_range=(40, 160), sigma: float = 0.5,
                  loop_shift: float = 3.0, lig_types=SMALL_TYPES) -> Ensemble:
    """A seeded ensemble: a base receptor of ``base_atoms`` atoms packed at the
    full workload's density (5,000 atoms per 45 A sphere); conformation i
    keeps its first N_i ~ U[n_range] atoms (never cutting an N/O from its
    hydrogen), moves every heavy atom by a Gaussian jitter of ``sigma`` A
    per axis (a hydrogen moves with its parent), then shifts 1-3 contiguous
    index segments of 20-200 atoms rigidly by up to ``loop_shift`` A.  Its
    box is cubic with an even npts ~ U[npts_range], centred on its
    centroid.  Maps: ``lig_types`` + e + d."""
    from . import Job

    density = 5000.0 / (4.0 / 3.0 * math.pi * 45.0 ** 3)
    radius = (base_atoms / (4.0 / 3.0 * math.pi * density)) ** (1.0 / 3.0)
    os.makedirs(outdir, exist_ok=True)
    coords, types, q = make_receptor(base_atoms, radius, seed)
    rec = "ensemble.pdbqt"
    write_pdbqt(os.path.join(outdir, rec), coords, types, q)
    centroid = read_pdbqt_coords(os.path.join(outdir, rec)).mean(axis=0)
    present = []
    for t in types:
        if t not in present:
            present.append(t)
    gpf = "ensemble.gpf"
    write_gpf(os.path.join(outdir, gpf), "ensemble", rec, present, list(lig_types), npts_range[0], centroid)
    cwd = os.getcwd()
    os.chdir(outdir)
    try:
        base = Job(gpf).prepare().state()
    finally:
        os.chdir(cwd)
    center = np.array([float("%.3f" % v) for v in centroid])
    absolute = base.coord + center
    is_h = np.asarray(base.hbond) == 2   # donor hydrogens (HD) of the generator
    parent = np.arange(base.n_atoms)
    for i in range(1, base.n_atoms):
        if is_h[i]:
            parent[i] = parent[i - 1]
    confs = []
    for i in range(n_conf):
        r = np.random.default_rng([seed, i])
        n = int(r.integers(n_range[0], n_range[1] + 1))
        n = min(n, base.n_atoms)
        while n < base.n_atoms and is_h[n]:
            n += 1
        c = absolute[:n].copy()
        jit = r.normal(0.0, sigma, size=(n, 3))
        c += jit[parent[:n]]
        for _ in range(int(r.integers(1, 4))):
            length = int(r.integers(20, 201))
            start = int(r.integers(0, max(1, n - length)))
            v = r.normal(size=3)
            v *= r.uniform(0.0, loop_shift) / np.linalg.norm(v)
            c[start:start + length] += v
        npts = 2 * int(r.integers(npts_range[0] // 2, npts_range[1] // 2 + 1))
        ctr = np.array([float("%.3f" % v) for v in c.mean(axis=0)])
        confs.append({"n_atoms": n, "coord": c, "npts": npts, "center": ctr})
    return Ensemble(base=base, center=center, confs=confs, lig_types=list(lig_types), rec_types=present)


def conformation_state(ens: Ensemble, i: int, bonds: bool = True):
    """PreparedState of conformation i: the base state's atoms cut to N_i,
    coordinates in its own grid frame, its box, and (``bonds``) its bond
    vectors from the engine's host precompute."""
    import dataclasses
    from . import bond_vectors

    b = ens.base
    cf = ens.confs[i]
    n = cf["n_atoms"]
    npts = cf["npts"]
    coord = cf["coord"] - cf["center"]
    st = dataclasses.replace(
        b, coord=coord, charge=b.charge[:n].copy(), vol=b.vol[:n].copy(), solpar=b.solpar[:n].copy(),
        type=b.type[:n].copy(), hbond=b.hbond[:n].copy(), rexp=np.zeros(n, np.int32), rvector=np.zeros((n, 3)),
        rvector2=np.zeros((n, 3)), disorder=np.zeros(n, np.uint8), n1=(npts + 1,) * 3, ne=(npts // 2,) * 3,
        _keep=[])
    if bonds:
        rexp, rv, rv2, dis = bond_vectors(st, st.coord, st.type, st.hbond)
        st = dataclasses.replace(st, rexp=rexp, rvector=rv, rvector2=rv2, disorder=dis, _keep=[])
    return st
