"""Ginsu window rules (CPU): integer-exact windows from shot geometry, clipping, the interior
box that the shared image receives, and rejection of windows outside the grid."""
import math

import numpy as np
import pytest

from wavefd_amd import ginsu


def test_window_for_shot_closed_form():
    big = (120, 200, 300)  # (nz, ny, nx)
    d = (10.0, 12.5, 12.5)
    w = ginsu.window_for_shot(big, d, (0.0, 0.0, 0.0), (20.0, 1000.0, 1500.0), [20.0, 20.0], [900.0, 1250.0],
                              [1400.0, 2100.0], aperture=(100.0, 200.0), pad=10)
    # y: min 900 - 100 = 800 -> floor(800/12.5) = 64 - 10 = 54; max 1250 + 100 = 1350 -> ceil(108) + 1 + 10 = 119
    # x: min 1400 - 200 = 1200 -> 96 - 10 = 86; max 2100 + 200 = 2300 -> 184 + 1 + 10 = 195
    assert w.origin == (0, 54, 86)
    assert w.size == (120, 119 - 54, 195 - 86)


def test_window_clips_and_covers():
    rng = np.random.default_rng(3)
    big = (80, 150, 170)
    d = (10.0, 10.0, 10.0)
    for _ in range(200):
        src = (rng.uniform(0, 790), rng.uniform(0, 1490), rng.uniform(0, 1690))
        nr = int(rng.integers(1, 20))
        ry = rng.uniform(0, 1490, nr)
        rx = rng.uniform(0, 1690, nr)
        ap = (rng.uniform(0, 300), rng.uniform(0, 300))
        pad = int(rng.integers(0, 15))
        w = ginsu.window_for_shot(big, d, (0.0, 0.0, 0.0), src, np.zeros(nr), ry, rx, aperture=ap, pad=pad)
        oz, oy, ox = w.origin
        sz, sy, sx = w.size
        assert 0 <= oy and oy + sy <= big[1] and 0 <= ox and ox + sx <= big[2]
        # every point (with aperture) lies inside the window's cells
        for yy in list(ry) + [src[1]]:
            assert oy * d[1] <= max(0.0, yy - ap[0]) + 1e-9
            assert min((big[1] - 1) * d[1], yy + ap[0]) <= (oy + sy - 1) * d[1] + 1e-9
        for xx in list(rx) + [src[2]]:
            assert ox * d[2] <= max(0.0, xx - ap[1]) + 1e-9
            assert min((big[2] - 1) * d[2], xx + ap[1]) <= (ox + sx - 1) * d[2] + 1e-9
        # the pad is kept where the grid allows
        lo = math.floor((min(list(ry) + [src[1]]) - ap[0]) / d[1]) - pad
        assert oy == max(0, lo)


def test_interior_box():
    big = (100, 100, 100)
    w = ginsu.Window(big, (0, 10, 0), (60, 50, 100))
    (z0, z1), (y0, y1), (x0, x1) = w.interior(8, freesurface=True)
    assert (z0, z1) == (0, 52)  # free-surface top kept; bottom sponge cut (not a grid face)
    assert (y0, y1) == (8, 42)  # both faces inside the grid: both sponges cut
    assert (x0, x1) == (0, 100)  # both faces are grid faces: nothing cut
    (z0, z1), _, _ = w.interior(8, freesurface=False)
    assert (z0, z1) == (0, 52)  # top is a grid face either way
    w2 = ginsu.Window(big, (5, 0, 0), (10, 10, 10))
    (z0, z1), _, _ = w2.interior(8)
    assert z0 <= z1 and 0 <= z0 and z1 <= 10


def test_window_outside_grid_rejected():
    with pytest.raises(ValueError):
        ginsu.Window((10, 10, 10), (0, 5, 0), (10, 6, 10))
    with pytest.raises(ValueError):
        ginsu.Window((10, 10, 10), (-1, 0, 0), (5, 5, 5))
    with pytest.raises(ValueError):
        ginsu.Window((10, 10, 10), (0, 0, 0), (0, 5, 5))
