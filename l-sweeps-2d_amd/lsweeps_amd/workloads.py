"""Synthetic problem definitions (media, frequencies, sources, boards).

The reference reads velocities from a CSV file (FunctionCSV, row 0 = top,
src/FunctionCSV.cpp:46-236).  These generators produce the same kind of grid
in memory; ``write_csv`` writes a file the reference's Input could read.
"""
import math

import numpy as np


def lens_velocity(n_cells, h, amp=0.4, width=0.02, cx=0.5, cy=0.5):
    """c(x,y) = 1 - amp*exp(-((x-cx)^2+(y-cy)^2)/width) sampled at the centres
    of n_cells x n_cells cells covering [-h/2, 1+h/2]^2 (cell centres on the
    grid nodes).  Returns (velocities float32 top-row-first, file domain)."""
    xs = -h / 2 + (np.arange(n_cells) + 0.5) * ((1 + h) / n_cells)
    ys = (1 + h / 2) - (np.arange(n_cells) + 0.5) * ((1 + h) / n_cells)
    X, Y = np.meshgrid(xs, ys)
    c = 1.0 - amp * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / width)
    return c.astype(np.float32), ((-h / 2, 1 + h / 2), (-h / 2, 1 + h / 2))


def write_csv(path, velocities):
    with open(path, "w") as f:
        for row in np.asarray(velocities, dtype=np.float32):
            f.write(",".join(f"{v:.9g}" for v in row) + "\n")


def config_lens_1600():
    """The 1600x1600 Gaussian-lens workload on an 8x8 board.

    NX = NY = 199 with 8 cells per axis gives n = 8*200-1 = 1599 intervals
    (h = 1/1599, 1600 nodes per axis).  omega gives 10 points per wavelength
    at the minimum velocity 0.6; MinWavelengthsInPML is chosen for 24 PML
    cells (the PML width uses the maximum velocity 1.0, src/Input.cpp:834-847).
    """
    n = 1599
    h = 1.0 / n
    cmin, cmax = 0.6, 1.0
    omega = 2 * math.pi * cmin / (10 * h)
    wl = 24 * h * omega / (2 * math.pi * cmax)
    vel, vdom = lens_velocity(n + 1, h)
    return dict(name="lens1600_8x8", domain=((0.0, 1.0), (0.0, 1.0)), ncells=(8, 8), nels=(199, 199),
                omega=omega, min_wl_in_pml=wl, velocity=vel, vel_domain=vdom, sources=[(0.5, 0.1)],
                tol=1e-8, max_iter=30, grid=(n + 1, n + 1))


def config_const_400():
    """The 400x400 constant-medium workload on a 4x4 board (h = 1/399,
    omega = 2*pi*40, 20 PML cells, source at the centre)."""
    return dict(name="const400_4x4", domain=((0.0, 1.0), (0.0, 1.0)), ncells=(4, 4), nels=(99, 99),
                omega=2 * math.pi * 40, min_wl_in_pml=2.0, velocity=None, vel_domain=None,
                sources=[(0.5, 0.5)], tol=1e-8, max_iter=30, grid=(400, 400))


CONFIGS = {"lens1600": config_lens_1600, "const400": config_const_400}
