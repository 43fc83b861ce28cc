"""Morphology of a phase field: the quantities long-horizon runs are compared on.

Dendritic growth amplifies rounding differences (the angle update of
src/Kobayashi.cpp:154-167 switches on |grad phi| > FLT_EPSILON), so two
correct arithmetic paths agree per cell only for short horizons; beyond that
they agree on shape:
  solid_fraction   fraction of cells with phi > 1/2 (kg_reduce computes it on device)
  radius_profile   distance from a nucleus centre to the phi = 1/2 front along
                   `n` directions (ray marching with linear interpolation)
  tip_radius       largest front distance (the dendrite tips)
  arm_count        peaks of the radius profile whose prominence is at least
                   `prominence` x (max - min radius) -- the number of primary arms
                   (side branches of grown dendrites stand out far less)
  tip_velocity     growth rate of tip_radius between two snapshots
device_summary computes the same quantities on the GPU (kg_reduce and
kg_front_profile, the latter the radius_profile algorithm in FP64).
"""
from __future__ import annotations

import numpy as np


def solid_fraction(phi: np.ndarray) -> float:
    return float((np.asarray(phi) > 0.5).mean())


def radius_profile(phi: np.ndarray, cx: float, cy: float, n: int = 360,
                   level: float = 0.5) -> np.ndarray:
    """Front distance (cells) from (cx, cy) along n equally spaced angles."""
    phi = np.asarray(phi, dtype=np.float64)
    ny, nx = phi.shape
    rmax = float(np.hypot(nx, ny))
    ang = 2 * np.pi * np.arange(n) / n
    steps = np.arange(0.0, rmax, 0.25)
    xs = cx + np.outer(np.cos(ang), steps)
    ys = cy + np.outer(np.sin(ang), steps)
    inside = (xs >= 0) & (xs <= nx - 1) & (ys >= 0) & (ys <= ny - 1)
    # bilinear sample
    x0 = np.clip(np.floor(xs).astype(int), 0, nx - 2)
    y0 = np.clip(np.floor(ys).astype(int), 0, ny - 2)
    fx = np.clip(xs - x0, 0, 1)
    fy = np.clip(ys - y0, 0, 1)
    v = (phi[y0, x0] * (1 - fx) * (1 - fy) + phi[y0, x0 + 1] * fx * (1 - fy)
         + phi[y0 + 1, x0] * (1 - fx) * fy + phi[y0 + 1, x0 + 1] * fx * fy)
    v = np.where(inside, v, 0.0)
    out = np.zeros(n)
    for k in range(n):
        below = np.nonzero(v[k] < level)[0]
        if len(below) == 0:
            out[k] = steps[-1]
            continue
        i = below[0]
        if i == 0:
            out[k] = 0.0
            continue
        a, b = v[k, i - 1], v[k, i]
        t = (a - level) / (a - b) if a != b else 0.0
        out[k] = steps[i - 1] + t * (steps[i] - steps[i - 1])
    return out


def tip_radius(phi: np.ndarray, cx: float, cy: float, n: int = 720) -> float:
    return float(radius_profile(phi, cx, cy, n).max())


def arm_count(phi: np.ndarray, cx: float, cy: float, n: int = 720,
              prominence: float = 0.3, min_relief: float = 0.15) -> int:
    """Number of circular peaks of the front radius whose height above the
    neighbouring minima exceeds `prominence` x (max - min radius); a profile
    whose relief (max - min) is below `min_relief` x mean radius has no arms."""
    return arms_of_profile(radius_profile(phi, cx, cy, n), prominence, min_relief)


def arms_of_profile(r: np.ndarray, prominence: float = 0.3, min_relief: float = 0.15) -> int:
    """arm_count of an already computed radius profile (e.g. the device's
    Solver.front_profile)."""
    r = np.asarray(r, dtype=np.float64)
    n = len(r)
    if r.max() <= 0 or r.max() - r.min() < min_relief * r.mean():
        return 0
    thr = prominence * (r.max() - r.min())
    # circular peak finding with a plateau-safe comparison
    peaks = []
    for k in range(n):
        if r[k] >= r[k - 1] and r[k] > r[(k + 1) % n]:
            peaks.append(k)
    count = 0
    for k in peaks:
        # walk both ways to the nearest point higher than r[k]; the lowest
        # point passed on either side bounds the peak's prominence
        best = np.inf
        for step in (1, -1):
            lo = r[k]
            for d in range(1, n):
                v = r[(k + step * d) % n]
                if v > r[k]:
                    break
                lo = min(lo, v)
            best = min(best, r[k] - lo) if lo < r[k] else best
        if best == np.inf:      # the global maximum: prominence against the global minimum
            best = r[k] - r.min()
        if best >= thr:
            count += 1
    return count


def tip_velocity(phi_a: np.ndarray, phi_b: np.ndarray, cx: float, cy: float,
                 dt_between: float, dx: float) -> float:
    """(tip radius change) x dx / time between the snapshots (length per time)."""
    return (tip_radius(phi_b, cx, cy) - tip_radius(phi_a, cx, cy)) * dx / dt_between


def summary(phi: np.ndarray, cx: float, cy: float) -> dict:
    return {"solid_fraction": solid_fraction(phi), "tip_radius": tip_radius(phi, cx, cy),
            "arm_count": arm_count(phi, cx, cy)}


def device_summary(solver, cx: float, cy: float, n: int = 720) -> dict:
    """summary() computed on the GPU: FP64 reductions (kg_reduce) and the
    front profile (kg_front_profile); no field crosses PCIe."""
    r = solver.front_profile(cx, cy, n)
    return {"solid_fraction": solver.stats()["solid_fraction"], "tip_radius": float(r.max()),
            "arm_count": arms_of_profile(r)}
