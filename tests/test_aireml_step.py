"""The trust-region step that replaces NLopt LD_MMA (varCompEst.cpp:1162-1214).

NLopt 2.7.1 ships only as a binary inside the reference, so its trajectory is not
reproducible here. What the step must be is fixed by the problem it solves:
  * unconstrained case (the reference's usual one: Delta starts at 1e20): the Newton
    step AI^-1 g, the unique minimiser of the convex quadratic;
  * binding constraints: a feasible KKT point, checked against scipy's SLSQP on the
    same problem (objective, box, signed trust region, positivity, covariance).
"""
import ctypes as C

import numpy as np
import pytest
from scipy.optimize import minimize

import globe_amd


def step(g, A, th, delta):
    L = globe_amd.load_library()
    L.globe_aireml_step.argtypes = [C.c_void_p] * 3 + [C.c_double, C.c_void_p]
    p = np.zeros(5)
    g = np.ascontiguousarray(g, dtype=np.float64)
    A = np.ascontiguousarray(A, dtype=np.float64)
    th = np.ascontiguousarray(th, dtype=np.float64)
    rc = L.globe_aireml_step(g.ctypes.data, A.ctypes.data, th.ctypes.data, delta, p.ctypes.data)
    return rc, p


def spd(rng, scale=1.0):
    M = rng.standard_normal((5, 5))
    return scale * (M @ M.T + 0.5 * np.eye(5))


def objective(p, g, A):
    return -(g @ p - 0.5 * p @ A @ p)


def constraints(th, A, delta):
    cons = [{"type": "ineq", "fun": lambda p: delta - np.diag(A) @ p}]
    for i in (0, 1, 2, 4):
        cons.append({"type": "ineq", "fun": lambda p, i=i: th[i] + p[i] - 1e-5})
    cons.append({"type": "ineq", "fun": lambda p: -((th[3] + p[3]) ** 2 - (th[2] + p[2]) * (th[4] + p[4]) + 1e-3)})
    return cons


def feasible(p, th, A, delta, tol=1e-8):
    ok = np.all(p >= -10 - tol) and np.all(p <= 10 + tol)
    ok &= np.diag(A) @ p <= delta + tol * (1 + abs(delta))
    ok &= all(th[i] + p[i] >= 1e-5 - tol for i in (0, 1, 2, 4))
    ok &= (th[3] + p[3]) ** 2 - (th[2] + p[2]) * (th[4] + p[4]) + 1e-3 <= tol
    return bool(ok)


def test_unconstrained_is_newton_step():
    rng = np.random.default_rng(0)
    A = spd(rng, 100.0)
    th = np.array([0.5, 0.3, 0.4, 0.05, 0.3])
    pn = rng.standard_normal(5) * 1e-3
    g = A @ pn
    rc, p = step(g, A, th, 1e20)
    assert rc == 0 and np.allclose(p, np.linalg.solve(A, g), rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("seed", range(12))
def test_constrained_step_matches_slsqp(seed):
    rng = np.random.default_rng(100 + seed)
    A = spd(rng, 50.0)
    th = np.array([0.5, 0.3, 0.3, 0.1, 0.2])
    # Newton steps large enough to violate positivity / covariance, some trust regions binding
    g = A @ (rng.standard_normal(5) * 0.6)
    delta = 1e20 if seed % 3 == 0 else float(abs(np.diag(A) @ np.linalg.solve(A, g)) * 0.3 + 0.5)
    rc, p = step(g, A, th, delta)
    assert rc == 0
    assert feasible(p, th, A, delta)
    best = None
    for x0 in [np.zeros(5), p, rng.standard_normal(5) * 0.01, np.linalg.solve(A, g) * 0.5]:
        r = minimize(objective, x0, args=(g, A), method="SLSQP", bounds=[(-10, 10)] * 5,
                     constraints=constraints(th, A, delta), options={"ftol": 1e-14, "maxiter": 1000})
        # SLSQP often reports a line-search failure at an optimum: judge by feasibility
        if feasible(r.x, th, A, delta, 1e-7) and (best is None or r.fun < best):
            best = r.fun
    assert best is not None
    # the step is at least as good as SLSQP's best local solution (to solver tolerance)
    assert objective(p, g, A) <= best + 1e-6 * (1 + abs(best))
