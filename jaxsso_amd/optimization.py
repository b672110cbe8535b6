"""Host optimisation driver around the GPU objective + gradient.

Mirrors the reference's ``Optimization`` class (JaxSSO/optimization.py:12-372):
the same constructor, setters and ``optimize(x_ini, log)``.  The objective is a
callable ``fun(x) -> (value, grad)``, e.g. a wrapper of
``SSO_model.value_grad_params``; each call is one compliance + gradient
evaluation on the GPU, and the driver itself is a few numpy operations per step.

Methods:
    'GD'    gradient descent, optionally with the gradient scaled by its largest
            absolute entry (JaxSSO/optimization.py:342-372)
    'SLSQP' needs the nlopt package (as in the reference); nlopt is not part of
            this build's environment, so choosing it raises ImportError
"""
from __future__ import annotations

import numpy as np


class Optimization:
    """Optimizer for ``n_x`` parameters (JaxSSO/optimization.py:12-36)."""

    def __init__(self, n_x, method="GD"):
        if method not in ("GD", "SLSQP"):
            raise ValueError(f"unknown method {method!r} ('GD' or 'SLSQP')")
        if method == "SLSQP":
            try:
                import nlopt  # noqa: F401
            except ImportError as e:
                raise ImportError("method 'SLSQP' needs the nlopt package, which is not installed") from e
            raise NotImplementedError("method 'SLSQP' is not provided by this build; use 'GD'")
        self.method = method
        self.n_x = int(n_x)
        self.if_eq_c = False
        self.if_ineq_c = False
        self.if_bounds = False
        self.if_maxiter = False
        self.if_ftol_rel = False
        self.if_ftol_abs = False
        self.GD_norm = False
        self.ln_rate = None
        self.history = []

    def set_objective(self, fun):
        """``fun(x) -> (value, grad)`` with grad of shape (n_x,)."""
        self.fun = fun

    def set_maxiter(self, maxiter):
        self.if_maxiter = True
        self.maxiter = int(maxiter)

    def set_step_size(self, ln_rate):
        if self.method == "GD":
            self.ln_rate = float(ln_rate)

    def set_GD_normalized(self, normalized=False):
        """Scale the gradient by max |gradient| before each step."""
        self.GD_norm = bool(normalized)

    def set_ftol_rel(self, ftol_rel):
        """Stop when |f_{i-1} - f_i| / |f_{i-1}| <= ftol_rel."""
        self.if_ftol_rel = True
        self.ftol_rel = float(ftol_rel)

    def set_ftol_abs(self, ftol_abs):
        """Stop when |f_{i-1} - f_i| <= ftol_abs."""
        self.if_ftol_abs = True
        self.ftol_abs = float(ftol_abs)

    def set_bounds(self, x_lb, x_ub):
        # the reference's gradient descent has no bound handling either
        raise TypeError(" Gradient descent cannot deal with bounds, change 'method'")

    def set_inequality_constraints(self, fc, grad_fc, d_fc, tol=0):
        raise TypeError("constraints need method 'SLSQP' (nlopt), which this build does not provide")

    def set_equality_constraints(self, fc, grad_fc, d_fc, tol=0):
        raise TypeError("constraints need method 'SLSQP' (nlopt), which this build does not provide")

    def optimize(self, x_ini, log=True):
        """Gradient descent from ``x_ini``.

        Step i evaluates f and its gradient at x_i, checks the ftol criteria
        against f_{i-1} (stopping there, at x_i), then moves
        x_{i+1} = x_i - step * g (g scaled by max |g| when normalised).  After
        maxiter steps the last moved parameters are returned, as in the
        reference.  The objective values are kept in ``self.history``."""
        if not self.if_maxiter:
            raise ValueError("set_maxiter() before optimize()")
        if self.ln_rate is None:
            raise ValueError("set_step_size() before optimize()")
        x = np.array(x_ini, dtype=float, copy=True)
        if x.shape != (self.n_x,):
            raise ValueError(f"x_ini has shape {x.shape}, expected ({self.n_x},)")
        f = []
        i = 0
        while i < self.maxiter:
            obj, grad = self.fun(x)
            obj = float(obj)
            g = np.asarray(grad, dtype=float)
            f.append(obj)
            if self.GD_norm:
                m = np.max(np.abs(g))
                if m > 0:
                    g = g / m
            if i >= 1:
                if self.if_ftol_rel and abs((f[i - 1] - f[i]) / f[i - 1]) <= self.ftol_rel:
                    print("Stopping criteria 'ftol_rel' met, end of optimization")
                    break
                if self.if_ftol_abs and abs(f[i - 1] - f[i]) <= self.ftol_abs:
                    print("Stopping criteria 'ftol_abs' met, end of optimization")
                    break
            x = x - self.ln_rate * g
            if log:
                print(f"Step {i}, objective = {f[i]}")
            i += 1
        if i == self.maxiter:
            print("Stopping criteria 'maxiter' met, end of optimization")
        self.history = f
        return x
