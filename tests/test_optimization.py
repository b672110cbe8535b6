"""Host-side optimisation driver and parameter maps (CPU)."""
import os

import numpy as np
import pytest

from jaxsso_amd.optimization import Optimization
from jaxsso_amd.parameterization import linear_filter, load_quad_mesh_csv, simp_chain, simp_modulus

MESH = os.path.join(os.path.dirname(__file__), "golden", "mannheim_quad")


def test_load_quad_mesh_csv():
    x, y, z, cnct, bc = load_quad_mesh_csv(MESH)
    assert x.shape == y.shape == z.shape == (510,)
    assert cnct.shape == (457, 4) and cnct.min() == 0 and cnct.max() == 509
    assert all(len(set(q)) == 4 for q in cnct.tolist())
    assert bc.shape == (98,) and len(set(bc.tolist())) == 98 and bc.max() < 510
    # every node is used by some quad
    assert np.array_equal(np.unique(cnct), np.arange(510))


def test_linear_filter_definition_and_invariants():
    x, y, _, _, _ = load_quad_mesh_csv(MESH)
    xs, ys = x - x.min(), y - y.min()
    R = 10.0
    B = linear_filter(xs, ys, R)
    Bd = B.toarray()
    D = np.hypot(xs[:, None] - xs[None, :], ys[:, None] - ys[None, :])
    W = np.where(D > R, 0.0, (R - D) / R)          # the notebook's hat weights
    assert np.abs(Bd - W / W.sum(1)[:, None]).max() < 1e-15
    assert np.allclose(Bd.sum(1), 1.0, rtol=0, atol=1e-15)
    assert np.all(Bd[D > R] == 0.0)
    # the unnormalised weights are symmetric; row i is scaled by 1 / sum_j w_ij
    w = Bd * W.sum(1)[:, None]
    assert np.abs(w - w.T).max() < 1e-13
    # a constant field is reproduced exactly
    assert np.allclose(B @ np.full(510, 3.25), 3.25, rtol=0, atol=1e-14)


def test_simp_chain_matches_finite_difference():
    rho = np.linspace(0.01, 1.0, 7)
    E0, p, h = 30e9, 3.0, 1e-6
    fd = (simp_modulus(rho + h, E0, p) - simp_modulus(rho - h, E0, p)) / (2 * h)
    dE = np.full_like(rho, 2.5)   # dC/dE
    assert np.allclose(simp_chain(dE, rho, E0, p), dE * fd, rtol=1e-8)


def _quadratic(a):
    return lambda x: (0.5 * float(np.sum((x - a) ** 2)), x - a)


def test_gradient_descent_steps_by_hand():
    a = np.array([1.0, -2.0, 0.5])
    opt = Optimization(3, method="GD")
    opt.set_objective(_quadratic(a))
    opt.set_maxiter(3)
    opt.set_step_size(0.5)
    x = opt.optimize(np.zeros(3), log=False)
    # plain steps: x_{i+1} = x_i - 0.5 (x_i - a): x_3 = a (1 - 0.5^3)
    assert np.allclose(x, a * (1 - 0.5 ** 3), rtol=0, atol=1e-15)
    assert len(opt.history) == 3
    # normalised steps move the largest-gradient component by exactly the step
    opt2 = Optimization(3, method="GD")
    opt2.set_objective(_quadratic(a))
    opt2.set_maxiter(1)
    opt2.set_step_size(0.1)
    opt2.set_GD_normalized(True)
    x2 = opt2.optimize(np.zeros(3), log=False)
    g = -a
    assert np.allclose(x2, -0.1 * g / np.abs(g).max(), rtol=0, atol=1e-16)


def test_gradient_descent_stopping_rules_and_errors():
    a = np.array([1.0, 1.0])
    opt = Optimization(2, method="GD")
    opt.set_objective(_quadratic(a))
    opt.set_maxiter(1000)
    opt.set_step_size(0.1)
    opt.set_ftol_rel(0.2)
    x = opt.optimize(np.zeros(2), log=False)
    # f_i = f_0 0.81^i: the relative decrease 0.19 <= 0.2 stops at the second evaluation, at x_1
    assert len(opt.history) == 2 and np.allclose(x, 0.1 * a, rtol=0, atol=1e-16)
    opt = Optimization(2, method="GD")
    opt.set_objective(_quadratic(a))
    opt.set_maxiter(1000)
    opt.set_step_size(0.5)
    opt.set_ftol_abs(1e-6)
    x = opt.optimize(np.zeros(2), log=False)
    f = opt.history
    assert abs(f[-2] - f[-1]) <= 1e-6 and len(f) < 1000
    with pytest.raises(TypeError):
        opt.set_bounds(np.zeros(2), np.ones(2))
    with pytest.raises(ImportError):
        Optimization(2, method="SLSQP")
    with pytest.raises(ValueError):
        Optimization(2).optimize(np.zeros(2))
