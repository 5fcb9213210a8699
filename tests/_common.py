"""Shared helpers for the tests: build small problems with the oracle and the
HIP library on the same inputs."""
import math

import numpy as np

from oracle.geometry import Checkerboard, ConstantMedium, CSVMedium


def lens_velocity(n, amp=0.4, width=0.02, cx=0.5, cy=0.5, dom=((0.0, 1.0), (0.0, 1.0))):
    """Gaussian-lens velocities on an n x n cell-centred grid (row 0 = top)."""
    xs = dom[0][0] + (np.arange(n) + 0.5) * (dom[0][1] - dom[0][0]) / n
    ys = dom[1][1] - (np.arange(n) + 0.5) * (dom[1][1] - dom[1][0]) / n
    X, Y = np.meshgrid(xs, ys)
    return (1.0 - amp * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / width)).astype(np.float32)


def make_pair(ncells, nels, omega, wl=2.0, velocity=None, precision=0, device=True,
              dom=((0.0, 1.0), (0.0, 1.0)), general=False):
    """(oracle Checkerboard, lsweeps_amd.DDCheckerboard) on the same inputs."""
    import lsweeps_amd as L
    if velocity is None:
        med = ConstantMedium()
    else:
        med = CSVMedium(velocity, [list(dom[0]), list(dom[1])])
    cb = Checkerboard([list(dom[0]), list(dom[1])], ncells, nels, omega, wl, med, general=general)
    p = L.dd_params(domain=dom, ncells=ncells, nels=nels, omega=omega, min_wl_in_pml=wl,
                    precision=precision, velocity=velocity, vel_domain=dom, general=general)
    dd = L.DDCheckerboard(p, device=device)
    return cb, dd


def relerr(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))
