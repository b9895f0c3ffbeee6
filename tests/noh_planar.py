"""Planar Noh (Noh 1987, JCP 72, 78-120) as the reference's Noh-planar-1d.py checks it.

The analytic solution restates tests/functional/Hydro/Noh/NohAnalyticSolution.py:30-65 for
nDim = 1; the error norms restate SimulationControl/Pnorm.py (gridpnorm: interval weights
clipped to [rmin, rmax] for L1/L2, a 0/1 slice for Linf); the driver restates the script's
SpheralController sequence (iterateIdealH at start-up, step(5), advance(goalTime))."""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "noh_planar_1d_lnorm_sph.json")


def reference_norms():
    with open(GOLDEN) as f:
        return json.load(f)


def solution(t, r, h0, gamma=5.0 / 3.0, rho0=1.0, v0=-1.0):
    """(v, u, rho, P, h) of the planar Noh problem at time t and positions r."""
    r = np.asarray(r, np.float64)
    rshock = -t / 3.0 * v0
    inside = np.abs(r) <= rshock
    rho_s = rho0 * 2.0 ** 2
    v = np.where(inside, 0.0, v0)
    u = np.where(inside, 0.5 * v0 ** 2, 0.0)
    rho = np.where(inside, rho_s, rho0)
    h = np.where(inside, h0 * (rho0 / rho_s) ** 1.0, h0)
    P = (gamma - 1.0) * u * rho
    return v, u, rho, P, h


def _grid_weights(pos, rmin, rmax):
    order = sorted(zip(pos.tolist(), range(len(pos))))
    n = len(order)
    w = np.zeros(n)
    for j in range(n):
        i = order[j][1]
        r0 = max(rmin, min(rmax, order[j][0] if j == 0 else 0.5 * (order[j - 1][0] + order[j][0])))
        r1 = max(rmin, min(rmax, order[j][0] if j == n - 1 else 0.5 * (order[j][0] + order[j + 1][0])))
        w[i] = r1 - r0
    return w


def gridpnorm(err, pos, p, rmin, rmax):
    err = np.abs(np.asarray(err, np.float64))
    pos = np.asarray(pos, np.float64)
    if p == "inf":
        w = ((pos >= rmin) & (pos <= rmax)).astype(np.float64)
        return float(np.linalg.norm(w * err, np.inf))
    w = _grid_weights(pos, rmin, rmax)
    return float(np.linalg.norm(w * err, p) / max(1e-30, w.sum()) ** (1.0 / p))


def norms(x, rho, P, v, eps, h, t, h0, rmin, rmax):
    va, ua, rhoa, Pa, ha = solution(t, x, h0)
    out = {}
    for name, data, ans in (("Mass density", rho, rhoa), ("Pressure", P, Pa), ("Velocity", v, va),
                            ("Spec Therm E", eps, ua), ("h", h, ha)):
        e = np.asarray(data) - ans
        out[name] = {k: gridpnorm(e, x, p, rmin, rmax) for k, p in (("L1", 1), ("L2", 2), ("Linf", "inf"))}
    return out


def within(ours, ref, tol):
    """np.allclose(L, ref, tol, tol) per norm, as the reference script tests."""
    bad = []
    for name, d in ref.items():
        for k, r in d.items():
            if not abs(ours[name][k] - r) <= tol + tol * abs(r):
                bad.append((name, k, ours[name][k], r))
    return bad


def drive(run, goal_time, steps_first=5, max_steps=100000):
    """SpheralController: step(5) (advance(1e40, 5)), then advance(goalTime)."""
    for _ in range(steps_first):
        run.step(1.0e40)
    n = 0
    while run.time()[0] < goal_time and n < max_steps:
        run.step(goal_time)
        n += 1
    return steps_first + n
