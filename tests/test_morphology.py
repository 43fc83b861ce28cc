"""Morphology metrics (crystalgrowth_amd/morphology.py) on CPU-oracle runs and
closed-form shapes."""
import numpy as np
import pytest

from crystalgrowth_amd import morphology as M


def test_disk_radius_and_no_arms():
    n = 101
    yy, xx = np.mgrid[0:n, 0:n]
    phi = (np.hypot(xx - 50, yy - 50) <= 20).astype(float)
    r = M.radius_profile(phi, 50, 50, 360)
    assert np.all(np.abs(r - 20.5) < 1.0)
    assert M.arm_count(phi, 50, 50) == 0
    assert M.solid_fraction(phi) == pytest.approx(np.pi * 20 ** 2 / n ** 2, rel=0.05)


@pytest.mark.parametrize("k", [3, 4, 6])
def test_star_arm_count(k):
    n = 201
    yy, xx = np.mgrid[0:n, 0:n]
    ang = np.arctan2(yy - 100, xx - 100)
    rad = 40 + 30 * np.cos(k * ang)
    phi = (np.hypot(xx - 100, yy - 100) <= rad).astype(float)
    assert M.arm_count(phi, 100, 100) == k
    assert M.tip_radius(phi, 100, 100) == pytest.approx(70, abs=1.5)


@pytest.mark.parametrize("j", [4, 6])
def test_oracle_dendrite_has_j_arms(O, j):
    """The reference model grows j primary arms (docs/images gallery: 6-fold)."""
    p = O.default_params(aniso=float(j), delta=0.05)
    s = O.run(O.init_plus(160, 160, np.float32), p, 1200)
    assert M.arm_count(s.phi, 80, 80) == j
    assert M.tip_radius(s.phi, 80, 80) > 25


def test_tip_velocity_of_growing_disk():
    """A disk grown from radius 20 to 30 cells in 0.1 time units at dx = 0.03
    moves its tip at 10 * 0.03 / 0.1 = 3 length units per time unit."""
    n = 101
    yy, xx = np.mgrid[0:n, 0:n]
    r = np.hypot(xx - 50, yy - 50)
    a = (r <= 20).astype(float)
    b = (r <= 30).astype(float)
    v = M.tip_velocity(a, b, 50, 50, dt_between=0.1, dx=0.03)
    assert v == pytest.approx(3.0, abs=0.05)
    # a shrinking front has a negative velocity; a still one none
    assert M.tip_velocity(b, a, 50, 50, 0.1, 0.03) == pytest.approx(-3.0, abs=0.05)
    assert M.tip_velocity(a, a, 50, 50, 0.1, 0.03) == 0.0


def test_tip_velocity_of_star_arms():
    """Tips of a 6-armed star whose arms lengthen by 12 cells."""
    n = 241
    yy, xx = np.mgrid[0:n, 0:n]
    ang = np.arctan2(yy - 120, xx - 120)
    rad = np.hypot(xx - 120, yy - 120)
    a = (rad <= 40 + 30 * np.cos(6 * ang)).astype(float)
    b = (rad <= 40 + 42 * np.cos(6 * ang)).astype(float)
    v = M.tip_velocity(a, b, 120, 120, dt_between=0.5, dx=0.03)
    assert v == pytest.approx(12 * 0.03 / 0.5, rel=0.08)
