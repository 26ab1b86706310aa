"""Fitting harness for the analytic map mode (PyTorch).

This module is the only place torch touches the engine: a
``torch.autograd.Function`` whose forward calls ``ag_analytic_forward`` and
whose backward calls the hand-written adjoint ``ag_analytic_adjoint``, a
trilinear pose scorer, and a weight fit with Adam.  Import torch before the
engine library is first loaded (this module does), so both use one HIP
runtime.

Ligand poses are scored the way docking reads AutoGrid maps: each ligand atom
of type t, charge q at position x adds map_t(x) + q e(x) + |q| d(x), the
maps interpolated trilinearly between the 8 grid points around x
(.fld layout: x fastest, main.cpp:322).
"""
from __future__ import annotations

import torch  # noqa: F401  (first: one HIP runtime for torch and the engine)

import numpy as np

from . import Analytic, Context, PreparedState


class AnalyticMaps(torch.autograd.Function):
    """maps = AnalyticMaps.apply(coord, charge, rij, eps, weights, engine)

    coord [N, 3], charge [N], rij / eps [M, T] (eps unweighted), weights [4]
    (vdW, H-bond, electrostatic, desolvation): float64 CPU tensors.  Returns
    the [n_maps, nz, ny, nx] float32 map set on the engine's GPU.  The
    backward pass runs the engine's adjoint with dL/dmaps."""

    @staticmethod
    def forward(ctx, coord, charge, rij, eps, weights, engine: Analytic):
        engine.set(coord.detach().cpu().numpy(), charge.detach().cpu().numpy(), rij.detach().cpu().numpy(),
                   eps.detach().cpu().numpy(), weights.detach().cpu().numpy())
        maps = torch.empty(engine.shape, dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        engine.forward(maps.data_ptr())
        ctx.engine = engine
        return maps

    @staticmethod
    def backward(ctx, dmaps):
        dmaps = dmaps.contiguous().to(torch.float32)
        torch.cuda.synchronize()
        g = ctx.engine.adjoint(dmaps.data_ptr())
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a))  # noqa: E731
        return t(g["coord"]), t(g["charge"]), t(g["rij"]), t(g["eps"]), t(g["weights"]), None


def make_poses(st: PreparedState, map_of_type: dict, n_poses: int = 256, seed: int = 1005,
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


class PoseScorer:
    """Scores of ligand poses against a map set: score = S @ maps.ravel(),
    S the sparse trilinear-interpolation matrix [n_poses, n_maps * P]."""

    def __init__(self, st: PreparedState, poses, device="cuda"):
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
                            if isinstance(m, (int, np.integer)) and m < 0:
                                continue
                            rows.append(np.full(pt.size, k))
                            cols.append(np.asarray(m) * P + pt)
                            vals.append(w * scale)
        r = np.concatenate(rows)
        c = np.concatenate(cols)
        v = np.concatenate(vals)
        self.n_poses = len(poses)
        self.shape = (self.n_poses, st.n_maps * P)
        idx = torch.from_numpy(np.stack([r, c]))
        self.S = torch.sparse_coo_tensor(idx, torch.from_numpy(v).to(torch.float32), self.shape,
                                         device=device).coalesce()
        self.S_cpu = (r, c, v)

    def __call__(self, maps: torch.Tensor) -> torch.Tensor:
        return torch.sparse.mm(self.S, maps.reshape(-1, 1)).reshape(-1)

    def cotangent(self, dscore: np.ndarray) -> np.ndarray:
        """dL/dmaps (flat) for dL/dscore, on the host."""
        r, c, v = self.S_cpu
        out = np.zeros(self.shape[1])
        np.add.at(out, c, v * dscore[r])
        return out


def fit_weights(ctx: Context, st: PreparedState, poses, scale_true, steps: int = 200, lr: float = 0.02):
    """Fit the four AD4 weights to the scores the maps give at
    weights0 * scale_true, starting from weights0, with Adam on log-weights.
    Returns a dict with the loss history, the fitted and the true weights
    and the time per step (forward + adjoint)."""
    import time
    eng = Analytic(ctx, st)
    scorer = PoseScorer(st, poses)
    coord = torch.from_numpy(np.ascontiguousarray(st.coord, np.float64))
    charge = torch.from_numpy(np.ascontiguousarray(st.charge, np.float64))
    rij = torch.from_numpy(eng.rij0.copy())
    eps = torch.from_numpy(eng.eps0.copy())
    w0 = torch.from_numpy(eng.weights0.copy())
    w_true = w0 * torch.as_tensor(scale_true, dtype=torch.float64)
    with torch.no_grad():
        target = scorer(AnalyticMaps.apply(coord, charge, rij, eps, w_true, eng)).detach()
    theta = torch.zeros(4, dtype=torch.float64, requires_grad=True)
    opt = torch.optim.Adam([theta], lr=lr)
    hist = []
    t0 = time.perf_counter()
    for _ in range(steps):
        opt.zero_grad()
        w = w0 * torch.exp(theta)
        s = scorer(AnalyticMaps.apply(coord, charge, rij, eps, w, eng))
        loss = torch.mean((s - target) ** 2)
        loss.backward()
        opt.step()
        hist.append(float(loss.detach().cpu()))
    t = (time.perf_counter() - t0) / steps
    w_fit = (w0 * torch.exp(theta.detach())).numpy()
    eng.close()
    return {"loss": hist, "w_fit": w_fit, "w_true": w_true.numpy(), "w0": w0.numpy(), "s_per_step": t}


def main(argv=None):
    """Analytic-mode fit on a synthetic workload; prints one JSON line.

    --workload small|analytic   the 1,000-atom small case, or the
                                5,000-atom receptor of the full workload on
                                a 65^3 box (npts 64) with maps A C HD N OA
    --check-grad                compare the autograd.Function gradients
                                with a direct adjoint call"""
    import argparse
    import json
    import os
    import tempfile
    import time

    from . import Job, synth

    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="analytic", choices=["small", "analytic"])
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--poses", type=int, default=256)
    ap.add_argument("--lr", type=float, default=0.02)
    ap.add_argument("--seed", type=int, default=1005)
    ap.add_argument("--check-grad", action="store_true")
    ap.add_argument("--workdir", default=None)
    a = ap.parse_args(argv)
    work = a.workdir or tempfile.mkdtemp(prefix="agfit_")
    info = synth.make_workload(a.workload, work)
    cwd = os.getcwd()
    os.chdir(work)
    try:
        st = Job(info["gpf"]).prepare().state()
    finally:
        os.chdir(cwd)
    lig = {t: i for i, t in enumerate(synth.WORKLOADS[a.workload].ligand_types) if t in ("C", "OA", "N", "HD", "A")}
    poses = make_poses(st, lig, n_poses=a.poses, seed=a.seed)
    rng = np.random.default_rng(a.seed)
    scale = rng.uniform(0.8, 1.2, size=4)
    out = {"workload": a.workload, "atoms": st.n_atoms, "points": st.points, "maps": st.n_maps,
           "poses": a.poses, "steps": a.steps, "scale_true": scale.tolist()}
    with Context(0) as ctx:
        if a.check_grad:
            eng = Analytic(ctx, st)
            scorer = PoseScorer(st, poses)
            coord = torch.from_numpy(np.ascontiguousarray(st.coord, np.float64)).requires_grad_()
            charge = torch.from_numpy(np.ascontiguousarray(st.charge, np.float64)).requires_grad_()
            rij = torch.from_numpy(eng.rij0.copy()).requires_grad_()
            eps = torch.from_numpy(eng.eps0.copy()).requires_grad_()
            w = torch.from_numpy(eng.weights0.copy()).requires_grad_()
            maps = AnalyticMaps.apply(coord, charge, rij, eps, w, eng)
            scorer(maps).sum().backward()
            dm = torch.zeros(eng.shape, dtype=torch.float32, device="cuda").reshape(-1)
            dm += torch.from_numpy(scorer.cotangent(np.ones(scorer.n_poses)).astype(np.float32)).cuda()
            torch.cuda.synchronize()
            g = eng.adjoint(dm.data_ptr())
            rel = 0.0
            for name, t in (("coord", coord), ("charge", charge), ("rij", rij), ("eps", eps), ("weights", w)):
                ref = g[name]
                den = max(float(np.max(np.abs(ref))), 1e-30)
                rel = max(rel, float(np.max(np.abs(t.grad.numpy() - ref))) / den)
            out["grad_match_max_rel"] = rel
            eng.close()
        t0 = time.perf_counter()
        res = fit_weights(ctx, st, poses, scale, steps=a.steps, lr=a.lr)
        out["fit_s"] = time.perf_counter() - t0
    out["loss_initial"] = res["loss"][0]
    out["loss_final"] = res["loss"][-1]
    out["w_fit"] = res["w_fit"].tolist()
    out["w_true"] = res["w_true"].tolist()
    out["w_rel_err_max"] = float(np.max(np.abs(res["w_fit"] / res["w_true"] - 1.0)))
    out["s_per_step"] = res["s_per_step"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
