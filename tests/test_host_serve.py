"""CPU-side checks of the serving C-ABI: registration (host geometry and the
device-bytes estimate), argument errors, and that a request whose problem
cannot be loaded comes back failed instead of hanging.  No GPU is used: on a
machine without one, loading a problem fails inside the worker, which is the
failure path under test."""
import math

import numpy as np
import pytest

import lsweeps_amd as L


def _params(nc=(2, 2), nel=(19, 19), omega=2 * math.pi * 4.1, precision=L.LSW_PREC_FP32):
    return L.dd_params(ncells=nc, nels=nel, omega=omega, min_wl_in_pml=2.0, precision=precision)


def _has_gpu():
    try:
        return L.device_count() > 0
    except L.LswError:
        return False


def test_register_reports_grid_of_host_geometry():
    srv = L.Server(budget_bytes=64e9, slots=4, workers=1)
    try:
        pid = srv.add_problem(_params())
        geo = L.DDCheckerboard(_params(), device=False)
        assert srv.grid(pid) == (geo.Gx, geo.Gy)
        pid2 = srv.add_problem(_params(nc=(3, 3), nel=(12, 12)))
        assert pid2 == pid + 1
    finally:
        srv.close()


def test_problem_larger_than_budget_is_rejected_at_registration():
    srv = L.Server(budget_bytes=1e6, slots=4, workers=1)
    try:
        with pytest.raises(L.LswError) as e:
            srv.add_problem(_params())
        assert e.value.code == 5  # CAPACITY
    finally:
        srv.close()


@pytest.mark.parametrize("kw", [dict(slots=0), dict(slots=65), dict(workers=0), dict(workers=17),
                                dict(budget_bytes=0.0), dict(precision=7)])
def test_bad_server_options(kw):
    args = dict(budget_bytes=1e9, slots=4, workers=1)
    args.update(kw)
    with pytest.raises(L.LswError) as e:
        L.Server(**args)
    assert e.value.code == 1  # ARG


def test_submit_argument_errors():
    srv = L.Server(budget_bytes=64e9, slots=4, workers=1)
    try:
        with pytest.raises(L.LswError):
            srv.submit(0, [(0.5, 0.5)], want_solution=False)   # no problem registered
        pid = srv.add_problem(_params())
        with pytest.raises(L.LswError):
            srv.submit(pid, np.zeros((0, 2)), want_solution=False)   # no sources
    finally:
        srv.close()


@pytest.mark.skipif(_has_gpu(), reason="exercises the load failure of a machine without a GPU")
def test_load_failure_fails_every_request_of_the_problem():
    srv = L.Server(budget_bytes=64e9, slots=2, workers=2)
    try:
        pid = srv.add_problem(_params())
        ids = [srv.submit(pid, [(0.5, 0.5), (0.25, 0.5), (0.75, 0.5)], want_solution=False) for _ in range(3)]
        res = [srv.wait(i) for i in ids]
        assert all(r["status"] != 0 for r in res)
        assert all(r["message"] for r in res)
        # later requests to the failed problem fail at once
        r = srv.wait(srv.submit(pid, [(0.5, 0.5)], want_solution=False))
        assert r["status"] == 3  # STATE
        st = srv.stats()
        assert st["requests"] == 4 and st["completed"] == 4 and st["loads"] == 0
        assert st["reserved_bytes"] == 0
    finally:
        srv.close()
