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

from . import Analytic, Context, PreparedState, Scorer
from .poses import cotangent, make_poses, splat


def _sync(device) -> None:
    """Order the engine's stream after torch's work on a GPU device."""
    if torch.device(device).type == "cuda":
        torch.cuda.synchronize()


class AnalyticMaps(torch.autograd.Function):
    """maps = AnalyticMaps.apply(coord, charge, rij, eps, weights, engine)

    coord [N, 3], charge [N], rij / eps [M, T] (eps unweighted), weights [4]
    (vdW, H-bond, electrostatic, desolvation): float64 CPU tensors.  Returns
    the [n_maps, nz, ny, nx] float32 map set on the engine's GPU.  The
    backward pass runs the engine's adjoint with dL/dmaps.

    The engine's adjoint linearises at the parameters and tape of its last
    forward.  Each forward keeps its own parameters and the engine's
    generation number; when the engine has been set or run again since (two
    applies before one backward), backward sets this forward's parameters
    back and re-runs the forward into scratch memory before the adjoint, so
    every graph gets the gradients of its own inputs."""

    @staticmethod
    def forward(ctx, coord, charge, rij, eps, weights, engine: Analytic):
        params = tuple(np.array(t.detach().cpu().numpy(), np.float64) for t in (coord, charge, rij, eps, weights))
        engine.set(*params)
        maps = torch.empty(engine.shape, dtype=torch.float32, device=engine.device)
        _sync(engine.device)
        engine.forward(maps.data_ptr())
        ctx.engine = engine
        ctx.params = params
        ctx.generation = engine.generation
        return maps

    @staticmethod
    def backward(ctx, dmaps):
        eng = ctx.engine
        dmaps = dmaps.contiguous().to(torch.float32)
        if eng.generation != ctx.generation:
            eng.set(*ctx.params)
            scratch = torch.empty(eng.shape, dtype=torch.float32, device=eng.device)
            _sync(eng.device)
            eng.forward(scratch.data_ptr())
        _sync(eng.device)
        g = eng.adjoint(dmaps.data_ptr())
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a))  # noqa: E731
        return t(g["coord"]), t(g["charge"]), t(g["rij"]), t(g["eps"]), t(g["weights"]), None


class PoseScores(torch.autograd.Function):
    """scores = PoseScores.apply(maps, scorer): the poses' trilinear scores
    from the engine's scoring kernel; the backward pass is the kernel's
    adjoint S^T dL/dscores."""

    @staticmethod
    def forward(ctx, maps, scorer: Scorer):
        maps = maps.contiguous()
        scores = torch.empty(scorer.n_poses, dtype=torch.float32, device=maps.device)
        torch.cuda.synchronize()
        scorer.score(maps.data_ptr(), scores.data_ptr())
        ctx.scorer = scorer
        return scores

    @staticmethod
    def backward(ctx, dscores):
        dscores = dscores.contiguous().to(torch.float32)
        dmaps = torch.empty(ctx.scorer.shape, dtype=torch.float32, device=dscores.device)
        torch.cuda.synchronize()
        ctx.scorer.adjoint(dscores.data_ptr(), dmaps.data_ptr())
        return dmaps, None


class PoseScorer:
    """Scores of ligand poses against a map set, on the GPU: each atom of
    type t, charge q at x adds map_t(x) + q e(x) + |q| d(x), trilinearly
    interpolated (ag_scorer_*)."""

    def __init__(self, ctx: Context, st: PreparedState, poses):
        self.scorer = Scorer(ctx, st, poses)
        self.n_poses = self.scorer.n_poses
        self.S_cpu = splat(st, poses)
        self.shape = (self.n_poses, st.n_maps * st.points)

    def __call__(self, maps: torch.Tensor) -> torch.Tensor:
        return PoseScores.apply(maps, self.scorer)

    def cotangent(self, dscore: np.ndarray) -> np.ndarray:
        """dL/dmaps (flat) for dL/dscore, on the host (FP64 reference)."""
        return cotangent(self.S_cpu, self.shape[1], dscore)

    def close(self):
        self.scorer.close()


def fit_weights(ctx: Context, st: PreparedState, poses, scale_true, steps: int = 200, lr: float = 0.02):
    """Fit the four AD4 weights to the scores the maps give at
    weights0 * scale_true, starting from weights0, with Adam on log-weights.

    The loss is the mean relative squared error ((s - t) / (1 + |t|))^2.
    Poses drawn uniformly in a box inside the receptor span scores from
    about -3 to 1e6 kcal/mol (steric clashes, pair energies capped at 1e5):
    under a plain squared error the clash poses fix the vdW and H-bond
    weights and leave the electrostatic and desolvation weights below the
    FP32 rounding of those scores (singular values of d score / d log w
    span 1.4e6 to 8 on the 5,000-atom, 65^3 workload), so they are not
    recovered.  Scaling each residual by its target gives every pose the
    same say.  Returns a dict with the loss history, the fitted and the
    true weights and the time per step (forward + adjoint)."""
    import time
    eng = Analytic(ctx, st)
    scorer = PoseScorer(ctx, st, poses)
    coord = torch.from_numpy(np.ascontiguousarray(st.coord, np.float64))
    charge = torch.from_numpy(np.ascontiguousarray(st.charge, np.float64))
    rij = torch.from_numpy(eng.rij0.copy())
    eps = torch.from_numpy(eng.eps0.copy())
    w0 = torch.from_numpy(eng.weights0.copy())
    w_true = w0 * torch.as_tensor(scale_true, dtype=torch.float64)
    with torch.no_grad():
        target = scorer(AnalyticMaps.apply(coord, charge, rij, eps, w_true, eng)).detach()
    inv = 1.0 / (1.0 + target.abs())
    theta = torch.zeros(4, dtype=torch.float64, requires_grad=True)
    opt = torch.optim.Adam([theta], lr=lr)
    hist = []
    t0 = time.perf_counter()
    for _ in range(steps):
        opt.zero_grad()
        w = w0 * torch.exp(theta)
        s = scorer(AnalyticMaps.apply(coord, charge, rij, eps, w, eng))
        loss = torch.mean(((s - target) * inv) ** 2)
        loss.backward()
        opt.step()
        hist.append(float(loss.detach().cpu()))
    t = (time.perf_counter() - t0) / steps
    w_fit = (w0 * torch.exp(theta.detach())).numpy()
    scorer.close()
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
            scorer = PoseScorer(ctx, st, poses)
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
            # the scoring kernel's scores against the FP64 host splat
            with torch.no_grad():
                s_dev = scorer(maps).double().cpu().numpy()
            r_, c_, v_ = scorer.S_cpu
            s_ref = np.zeros(scorer.n_poses)
            np.add.at(s_ref, r_, v_ * maps.detach().reshape(-1).double().cpu().numpy()[c_])
            out["score_max_rel"] = float(np.max(np.abs(s_dev - s_ref) / np.maximum(1.0, np.abs(s_ref))))
            scorer.close()
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
