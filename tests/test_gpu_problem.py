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


def test_problem_run_csv_file_medium(tmp_path):
    """A reference-format input whose squared slowness is Type = "CSV": the
    velocities written by workloads.write_csv on a 4 x 16 file grid over the
    unit square (mH = (1/16, 1/4): the clamp from above on x sits at
    1 - 1/8, so points past x = 0.875 read column 14, never 15).  The GPU
    Problem (file read with stof, reference GMRES) against the oracle's
    FunctionCSV read of the same file."""
    from lsweeps_amd.workloads import write_csv
    from oracle.geometry import CSVMedium
    rng = np.random.default_rng(9)
    vel = (1.0 + rng.random((4, 16))).astype(np.float32)
    write_csv(str(tmp_path / "vel.csv"), vel)
    omega = 2 * math.pi * 3.0
    cfg = tmp_path / "csv.cfg"
    cfg.write_text(f"""
Problems = ( {{ DomainDecomposition = "CheckerboardDD"; DomainDecompositionID = 0;
               RHS = "src"; Solver = "gmres"; Output = ""; }} );
Domains = {{ unit = {{ Dim = 2; ax = 0.0; bx = 1.0; ay = 0.0; by = 1.0; }}; }};
SquaredSlownesses = {{ m = {{ Type = "CSV"; Filename = "{tmp_path / 'vel.csv'}"; Dim = 2; Domain = "unit"; }}; }};
DomainDecompositions = {{ CheckerboardDD = ( {{
    Dim = 2; Domain = "unit"; PDE = "Helmholtz"; Omega = {omega!r};
    MinWavelengthsInPML = 2.0; SquaredSlowness = "m"; Order = 2;
    NX = 19; NY = 19; NCellsX = 2; NCellsY = 2; Solver = "HIP"; }} ); }};
RHSs = {{ src = ( {{ Dim = 2; x = 0.6; y = 0.4; }} ); }};
GlobalSolvers = {{ gmres = {{ Type = "GMRES"; Tol = 1.0e-8; MaxIter = 30; Verbose = false; }}; }};
""")
    r = L.problem_run(str(cfg), 0)
    med = CSVMedium.from_file(str(tmp_path / "vel.csv"), [[0.0, 1.0], [0.0, 1.0]])
    assert med.mH == [1.0 / 16, 0.25]
    cb = Checkerboard([[0.0, 1.0], [0.0, 1.0]], (2, 2), (19, 19), omega, 2.0, med)
    cb.factorize()
    A = cb.global_operator()
    xs, _ = cb.gcell.dof_points()
    assert np.any(xs > 0.9375)  # DOFs the clamp maps from column 15 to 14
    b = cb.global_rhs([(0.6, 0.4)])
    xo, log = gmres_reference(lambda v: A @ v, Preconditioner(cb).apply_global, b, tol=1e-8, max_iter=30)
    assert r["dofs"] == cb.N
    assert abs(r["iterations"] - (len(log) - 1)) <= 1
    assert relerr(r["x"], xo) < 1e-6
    assert np.linalg.norm(A @ r["x"] - b) / np.linalg.norm(b) < 1e-6


def test_dd_from_config_file_equals_dd_from_parameters():
    """lsw_dd_create_config builds the DD the input file describes: the same
    operator and preconditioner as the DD built from the same parameters, and
    the reference GMRES on it gives Problem's solution."""
    path = os.path.join(GOLD, "config_const_3x3.cfg")
    spec = L.config_problem(path, 0)
    a = L.DDCheckerboard(None, config=(path, 0))
    b = L.DDCheckerboard(L.dd_params(domain=spec["domain"], ncells=spec["ncells"], nels=spec["nels"],
                                     omega=spec["omega"], min_wl_in_pml=spec["min_wl_in_pml"]))
    assert (a.Gx, a.Gy, a.ncells) == (b.Gx, b.Gy, b.ncells)
    a.factorize()
    b.factorize()
    rng = np.random.default_rng(2)
    v = rng.standard_normal(a.N) + 1j * rng.standard_normal(a.N)
    assert np.array_equal(a.apply_a(v), b.apply_a(v))
    assert np.array_equal(a.apply_precond(v), b.apply_precond(v))
    rhs = a.assemble_rhs(spec["sources"])
    x, info = a.gmres(rhs, mode=0, tol=spec["tol"], max_iter=spec["max_iter"])
    r = L.problem_run(path, 0)
    assert relerr(x, r["x"]) < 1e-13 and info["iterations"] == r["iterations"]
