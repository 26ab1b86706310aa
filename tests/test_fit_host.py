"""The torch.autograd.Function of the fitting harness, on the CPU.

The engine's adjoint linearises at the parameters of its last forward, so a
graph whose forward is older than the engine's current state must be
re-linearised before its backward.  A stand-in engine with the same
set / forward / adjoint / generation protocol, whose gradient depends on
the parameters, shows whether each backward gets its own forward's
gradients: two applies with different weights, then the backward of the
first, against a fresh single apply.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from autogrid_amd.fit import AnalyticMaps  # noqa: E402


class FakeEngine:
    """maps = w0^2 * A + (sum coord) * B + w2 * q . C on a tiny map set; the
    adjoint returns the gradient at the parameters of the last forward."""

    def __init__(self, n_atoms=3, shape=(2, 2, 2, 2), seed=0):
        rng = np.random.default_rng(seed)
        self.shape = shape
        self.size = int(np.prod(shape))
        self.n_atoms = n_atoms
        self.A = rng.normal(size=self.size)
        self.B = rng.normal(size=self.size)
        self.Cq = rng.normal(size=(n_atoms, self.size))
        self.device = "cpu"
        self.generation = 0
        self.cur = None
        self.lin = None

    def set(self, coord, charge, rij, eps, weights):
        self.cur = [np.array(a, np.float64) for a in (coord, charge, rij, eps, weights)]
        self.generation += 1

    def _maps(self, p):
        coord, charge, _, _, w = p
        return w[0] ** 2 * self.A + coord.sum() * self.B + w[2] * charge @ self.Cq

    def forward(self, ptr):
        out = np.ctypeslib.as_array(C.cast(ptr, C.POINTER(C.c_float)), shape=(self.size,))
        out[:] = self._maps(self.cur)
        self.lin = [a.copy() for a in self.cur]
        self.generation += 1

    def adjoint(self, ptr):
        g = np.ctypeslib.as_array(C.cast(ptr, C.POINTER(C.c_float)), shape=(self.size,)).astype(np.float64)
        coord, charge, rij, eps, w = self.lin
        gw = np.zeros(4)
        gw[0] = 2 * w[0] * (g @ self.A)
        gw[2] = charge @ self.Cq @ g
        return {"coord": np.full(coord.shape, g @ self.B), "charge": w[2] * (self.Cq @ g),
                "rij": np.zeros(rij.shape), "eps": np.zeros(eps.shape), "weights": gw}


def _inputs(eng, w):
    coord = torch.arange(3 * eng.n_atoms, dtype=torch.float64).reshape(eng.n_atoms, 3) * 0.1
    charge = torch.tensor([0.3, -0.2, 0.1], dtype=torch.float64)
    rij = torch.ones(2, 2, dtype=torch.float64)
    eps = torch.ones(2, 2, dtype=torch.float64)
    weights = torch.tensor(w, dtype=torch.float64, requires_grad=True)
    return coord, charge, rij, eps, weights


def _grad(eng, w, cot):
    args = _inputs(eng, w)
    maps = AnalyticMaps.apply(*args, eng)
    (maps * cot).sum().backward()
    return args[4].grad.clone()


def test_backward_uses_its_own_forward():
    eng = FakeEngine()
    cot = torch.from_numpy(np.random.default_rng(1).normal(size=eng.shape)).float()
    w1, w2 = [0.8, 1.0, 1.3, 1.0], [1.7, 1.0, 0.4, 1.0]
    ref1 = _grad(FakeEngine(), w1, cot)
    ref2 = _grad(FakeEngine(), w2, cot)
    assert not torch.allclose(ref1, ref2)
    a1 = _inputs(eng, w1)
    m1 = AnalyticMaps.apply(*a1, eng)
    a2 = _inputs(eng, w2)
    m2 = AnalyticMaps.apply(*a2, eng)
    # backward of the first graph after the second forward
    (m1 * cot).sum().backward()
    assert torch.allclose(a1[4].grad, ref1, rtol=1e-6, atol=1e-9)
    (m2 * cot).sum().backward()
    assert torch.allclose(a2[4].grad, ref2, rtol=1e-6, atol=1e-9)


def test_backward_without_intervening_forward_reuses_tape():
    eng = FakeEngine()
    cot = torch.ones(eng.shape)
    a = _inputs(eng, [1.1, 1.0, 0.9, 1.0])
    m = AnalyticMaps.apply(*a, eng)
    gen = eng.generation
    m.sum().backward()
    # no re-linearisation needed: the engine still holds this forward
    assert eng.generation == gen
