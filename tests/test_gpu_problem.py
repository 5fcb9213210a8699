"""Problem::read + Problem::solve from a reference-format input file on the
GPU (src/Problem.cpp:70-158): the solution matches the oracle's reference
GMRES on the same inputs, and the VTK output is written when the problem
names one."""
import math
import os

import numpy as np
import pytest

import lsweeps_amd as L
from oracle.geometry import Checkerboard, ConstantMedium
from oracle.gmres import gmres_reference
from oracle.sweeps import Preconditioner
from tests._common import relerr

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _oracle_3x3():
    omega = 25.761059759436304
    cb = Checkerboard([[0.0, 1.0], [0.0, 1.0]], (3, 3), (12, 12), omega, 2.0, ConstantMedium())
    cb.factorize()
    A = cb.global_operator()
    b = cb.global_rhs([(0.5, 0.5)])
    return gmres_reference(lambda v: A @ v, Preconditioner(cb).apply_global, b, tol=1e-8, max_iter=30)


def test_problem_run_matches_oracle():
    r = L.problem_run(os.path.join(GOLD, "config_const_3x3.cfg"), 0)
    xo, log = _oracle_3x3()
    assert abs(r["iterations"] - (len(log) - 1)) <= 1
    assert relerr(r["x"], xo) < 1e-6
    assert r["factorize_s"] > 0 and r["gmres_s"] > 0


def test_problem_run_writes_vtk(tmp_path):
    txt = open(os.path.join(GOLD, "config_const_3x3.cfg")).read()
    txt = txt.replace('Output = "";', 'Output = "vtk";')
    txt = txt.replace('Filename = "out";', f'Filename = "{tmp_path / "wave"}";')
    cfg = tmp_path / "p.cfg"
    cfg.write_text(txt)
    r = L.problem_run(str(cfg), 0)
    assert (tmp_path / "wave.vtk").exists()
    raw = open(tmp_path / "wave.lswgrid", "rb").read()
    assert raw[:8] == b"LSWGRID1"
    gx, gy = np.frombuffer(raw[8:16], dtype="<i4")
    x = np.frombuffer(raw[16 + 8 * (gx + gy):], dtype="<c16")
    assert np.array_equal(x, r["x"])
