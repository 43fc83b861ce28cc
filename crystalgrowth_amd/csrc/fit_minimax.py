"""Fit the polynomial coefficients of kg_math.h (FP32 atan, sin, cos).

    python crystalgrowth_amd/csrc/fit_minimax.py

Method: near-minimax polynomials by iteratively reweighted least squares on a
dense grid (Lawson's algorithm), in float64, then rounded to float32:
  atan(q)  ~ q * A(q^2),          q in [0, 1]         (relative error)
  sin(r)   ~ r + r^3 * S(r^2),    r in [-pi/4, pi/4]  (relative error)
  cos(r)   ~ 1 + r^2 * C(r^2),    r in [-pi/4, pi/4]  (absolute error)
Cody-Waite constants split pi/2 into three floats with the FMA-based
reduction r = x - k*P1 - k*P2 - k*P3 (each step one fma).
Prints the coefficients and the float32-evaluated maximum errors.
"""
import numpy as np


def lawson(basis, target, weight, iters=200):
    """Weighted minimax fit target ~ basis @ c with error measured as
    (basis @ c - target) * weight."""
    n = len(target)
    w = np.full(n, 1.0 / n)
    for _ in range(iters):
        sw = np.sqrt(w)
        c, *_ = np.linalg.lstsq(basis * (sw * weight)[:, None], target * sw * weight, rcond=None)
        e = np.abs(basis @ c - target) * weight
        w = w * e
        w /= w.sum()
    return c


def fit_atan(deg=9):
    q = np.cos(np.linspace(0, np.pi / 2, 20001)) ** 1  # dense, clustered at both ends
    q = np.unique(np.concatenate([q, np.linspace(0, 1, 20001)]))[1:]
    t = np.arctan(q)
    basis = np.stack([q * q ** (2 * k) for k in range(deg)], 1)
    return lawson(basis, t, 1.0 / t)


def fit_sin(deg=4):
    r = np.linspace(1e-6, np.pi / 4, 20001)
    target = (np.sin(r) - r) / r ** 3
    basis = np.stack([r ** (2 * k) for k in range(deg)], 1)
    # relative error of sin: (r^3 S - (sin - r)) / sin
    return lawson(basis, target, r ** 3 / np.sin(r))


def fit_cos(deg=5):
    r = np.linspace(1e-6, np.pi / 4, 20001)
    target = (np.cos(r) - 1) / r ** 2
    basis = np.stack([r ** (2 * k) for k in range(deg)], 1)
    return lawson(basis, target, r ** 2)


def horner32(c, x):
    acc = np.float32(c[-1])
    for v in c[-2::-1]:
        acc = np.float32(acc * x + np.float32(v))
    return acc


def main():
    f32 = np.float32
    ca = fit_atan().astype(f32)
    cs = fit_sin().astype(f32)
    cc = fit_cos().astype(f32)
    q = np.linspace(0, 1, 200001).astype(f32)
    qa = (q * horner32(ca, q * q)).astype(np.float64)
    ea = np.max(np.abs(qa[1:] - np.arctan(q[1:].astype(float))) / np.arctan(q[1:].astype(float)))
    r = np.linspace(-np.pi / 4, np.pi / 4, 200001).astype(f32)
    r2 = r * r
    s = (r + r * r2 * horner32(cs, r2)).astype(float)
    c = (f32(1) + r2 * horner32(cc, r2)).astype(float)
    nz = np.abs(r) > 1e-3
    es = np.max(np.abs(s - np.sin(r.astype(float)))[nz] / np.abs(np.sin(r.astype(float)))[nz])
    ec = np.max(np.abs(c - np.cos(r.astype(float))))
    half_pi = np.pi / 2
    p1 = f32(half_pi)
    p2 = f32(half_pi - float(p1))
    p3 = f32(half_pi - float(p1) - float(p2))
    print("atan A[k] (q * sum A[k] q^(2k)):", ", ".join(f"{v:.9e}f" for v in ca))
    print("sin  S[k] (r + r^3 sum S[k] r^(2k)):", ", ".join(f"{v:.9e}f" for v in cs))
    print("cos  C[k] (1 + r^2 sum C[k] r^(2k)):", ", ".join(f"{v:.9e}f" for v in cc))
    print(f"pi/2 split: P1={p1:.9e}f P2={p2:.9e}f P3={p3:.9e}f")
    print(f"max rel err atan {ea:.3g}, max rel err sin {es:.3g}, max abs err cos {ec:.3g} "
          f"(float32 evaluation; 2^-24 = {2.0**-24:.3g})")


if __name__ == "__main__":
    main()
