"""CPU restatement of the reference's L-BFGS geometry optimiser
(src/optimizer.rs), over the oracle's energy + analytic gradient.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the device
optimiser.  The product never imports it.

Pinned by tests/test_optimizer_oracle.py:
  lbfgs_direction   against the explicit BFGS inverse-Hessian recursion
                    H <- (I - rho s y^T) H (I - rho y s^T) + rho s s^T from
                    H0 = gamma I (the textbook identity the two-loop recursion
                    evaluates), and H y_last = s_last (secant condition)
  limit_atom_step   closed form (largest per-atom norm equals the cap)
  optimize          exact minimiser of convex quadratics; the reference's own
                    stretched-H2 test (tests/optimizer.rs:5-22)
"""
import math

import numpy as np

import oracle as _oracle

C1 = 1.0e-4            # optimizer.rs:234
LINE_SEARCH_TRIALS = 24  # optimizer.rs:236


def dot(a, b):
    """Left-to-right sum, as the reference's iterator sum (optimizer.rs:325-327)."""
    s = 0.0
    for x, y in zip(a, b):
        s += x * y
    return s


def lbfgs_direction(g, s_hist, y_hist, rho_hist):
    """Two-loop recursion (optimizer.rs:259-287)."""
    if not s_hist:
        return [-v for v in g]
    q = list(g)
    m = len(s_hist)
    alpha = [0.0] * m
    for i in reversed(range(m)):
        alpha[i] = rho_hist[i] * dot(s_hist[i], q)
        q = [qk + (-alpha[i]) * yk for qk, yk in zip(q, y_hist[i])]
    yy = dot(y_hist[-1], y_hist[-1])
    ys = dot(y_hist[-1], s_hist[-1])
    gamma = ys / yy if yy > 1.0e-18 else 1.0
    q = [v * gamma for v in q]
    for i in range(m):
        beta = rho_hist[i] * dot(y_hist[i], q)
        q = [qk + (alpha[i] - beta) * sk for qk, sk in zip(q, s_hist[i])]
    return [-v for v in q]


def limit_atom_step(d, max_atom_step):
    """Per-atom step cap (optimizer.rs:289-305)."""
    if max_atom_step <= 0.0:
        return list(d)
    mx = 0.0
    for k in range(0, len(d) - 2, 3):
        n = math.sqrt(d[k] * d[k] + d[k + 1] * d[k + 1] + d[k + 2] * d[k + 2])
        mx = n if n > mx else mx
    if mx > max_atom_step:
        scale = max_atom_step / mx
        return [v * scale for v in d]
    return list(d)


def max_gradient(g):
    """Largest Cartesian component (gradient.rs:227-230)."""
    m = 0.0
    for v in g:
        m = max(m, abs(v))
    return m


def _finite(v):
    return all(math.isfinite(x) for x in v)


class EvaluationError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(message)
        self.status = status


def optimize(evaluate, x0, max_iterations=250, gradient_tolerance=1.0e-4, step_tolerance=1.0e-7, history=12,
             initial_step=1.0, max_atom_step=0.30):
    """optimize_geometry (optimizer.rs:78-185) over evaluate(x) -> (energy,
    gradient list); evaluate raises EvaluationError where the reference's
    evaluate returns Err.  Returns a dict of the result fields."""
    x = [float(v) for v in x0]
    e, g = evaluate(x)
    g = list(g)
    maxg = max_gradient(g)
    trajectory = [(0, e, maxg, 0.0, list(x))]
    history_len = max(history, 1)
    s_hist, y_hist, rho_hist = [], [], []
    converged = maxg <= gradient_tolerance
    iterations = 0
    evaluations = 1
    while not converged and iterations < max_iterations:
        iterations += 1
        d = lbfgs_direction(g, s_hist, y_hist, rho_hist)
        if dot(d, g) >= 0.0 or not _finite(d):
            d = [-v for v in g]
        d = limit_atom_step(d, max_atom_step)
        dd = dot(g, d)
        # line search (optimizer.rs:224-257)
        alpha = max(initial_step, 1.0e-8)
        last_error = None
        accepted = None
        for _ in range(LINE_SEARCH_TRIALS):
            trial = [xk + alpha * dk for xk, dk in zip(x, d)]
            try:
                evaluations += 1
                et, gt = evaluate(trial)
                if et <= e + C1 * alpha * dd:
                    accepted = (trial, et, list(gt))
                    break
            except EvaluationError as err:
                last_error = err
            alpha *= 0.5
        if accepted is None:
            if last_error is not None:
                raise last_error
            raise EvaluationError(3, "L-BFGS line search failed to find a downhill step")
        trial, et, gt = accepted
        s = [a - b for a, b in zip(trial, x)]
        step_norm = math.sqrt(dot(s, s))
        y = [a - b for a, b in zip(gt, g)]
        ys = dot(y, s)
        if ys > 1.0e-12 and _finite(s) and _finite(y):
            if len(s_hist) == history_len:
                s_hist.pop(0)
                y_hist.pop(0)
                rho_hist.pop(0)
            s_hist.append(s)
            y_hist.append(y)
            rho_hist.append(1.0 / ys)
        x, e, g = trial, et, gt
        maxg = max_gradient(g)
        trajectory.append((iterations, e, maxg, step_norm, list(x)))
        converged = maxg <= gradient_tolerance or step_norm <= step_tolerance
    return {"energy": e, "positions": np.array(x).reshape(-1, 3), "gradient": np.array(g).reshape(-1, 3),
            "iterations": iterations, "converged": converged, "max_gradient": maxg, "trajectory": trajectory,
            "evaluations": evaluations}


def optimize_gfn1(numbers, positions_bohr, charge=0.0, scc_options=None, **options):
    """optimize_geometry of one system over the oracle's total free energy and
    analytic gradient (gradient.rs:56-243)."""
    scc = dict(scc_options or {})
    scc["charge"] = charge

    def evaluate(x):
        r = _oracle.run(numbers, np.array(x).reshape(-1, 3), **scc)
        if r["status"] != 0:
            raise EvaluationError(r["status"], r["message"])
        return r["energy"], r["gradient"].reshape(-1).tolist()

    return optimize(evaluate, np.asarray(positions_bohr, dtype=np.float64).reshape(-1).tolist(), **options)
