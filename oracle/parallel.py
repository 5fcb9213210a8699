"""Oracle: the local solves of a checkerboard on P worker processes.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  Used by the CPU baseline
(oracle/cpu_run.py) to time the restatement the way the reference runs:
one process per group of cells, each holding its cells' factors
(DDCheckerboard::factorize, src/DDCheckerboard.cpp:648-665, runs on every
rank at once), and the local solves of one anti-diagonal of a sweep running
at the same time (the reference's ranks do them concurrently).

Cells go to workers round-robin by cell index (assign="cells") or by board
row (assign="rows": the reference's row-to-rank mapping,
src/DDCheckerboard.cpp:126-177).  Right-hand sides and solutions pass through
one shared-memory buffer with a slot per cell; a worker solves a slot in
place.  The pool forks, so it must be created by a process that has not
initialised a GPU runtime.
"""
import multiprocessing as mp
import time
from multiprocessing import shared_memory

import numpy as np


def _worker(conn, cells, owned, shm_name, offs, total):
    import scipy.sparse.linalg as spla
    shm = shared_memory.SharedMemory(name=shm_name)
    buf = np.ndarray((total,), dtype=np.complex128, buffer=shm.buf)
    try:
        t0 = time.perf_counter()
        lus = {i: spla.splu(cells[i].A.tocsc(), permc_spec="COLAMD") for i in owned}
        conn.send(("factorized", time.perf_counter() - t0))
        while True:
            msg = conn.recv()
            if msg is None:
                break
            for i in msg:
                v = buf[offs[i]:offs[i] + cells[i].size]
                v[:] = lus[i].solve(v)
            conn.send(len(msg))
    finally:
        del buf
        shm.close()


class CellPool:
    """solve_many([(cell index, rhs), ...]) on P processes."""

    def __init__(self, cb, procs, assign="cells"):
        self.cb = cb
        n = len(cb.cells)
        for c in cb.cells:
            if c.A is None:
                c.assemble()
        self.procs = max(1, min(int(procs), n))
        key = (lambda i: cb.cells[i].I) if assign == "rows" else (lambda i: i)
        self.owner = [key(i) % self.procs for i in range(n)]
        self.offs = np.concatenate([[0], np.cumsum([c.size for c in cb.cells])]).astype(np.int64)
        total = int(self.offs[-1])
        self.shm = shared_memory.SharedMemory(create=True, size=total * 16)
        self.buf = np.ndarray((total,), dtype=np.complex128, buffer=self.shm.buf)
        ctx = mp.get_context("fork")
        self.conns, self.ps = [], []
        t0 = time.perf_counter()
        for w in range(self.procs):
            a, b = ctx.Pipe()
            owned = [i for i in range(n) if self.owner[i] == w]
            p = ctx.Process(target=_worker, args=(b, cb.cells, owned, self.shm.name, self.offs, total), daemon=True)
            p.start()
            self.conns.append(a)
            self.ps.append(p)
        self.worker_factor_s = [c.recv()[1] for c in self.conns]
        self.factor_s = time.perf_counter() - t0   # wall time until every rank has its factors

    def solve_many(self, reqs):
        jobs = [[] for _ in range(self.procs)]
        for i, r in reqs:
            self.buf[self.offs[i]:self.offs[i + 1]] = r
            jobs[self.owner[i]].append(i)
        for w, js in enumerate(jobs):
            if js:
                self.conns[w].send(js)
        for w, js in enumerate(jobs):
            if js:
                self.conns[w].recv()
        return [self.buf[self.offs[i]:self.offs[i + 1]].copy() for i, _ in reqs]

    def close(self):
        for c in self.conns:
            try:
                c.send(None)
            except (BrokenPipeError, OSError):
                pass
        for p in self.ps:
            p.join(timeout=30)
        del self.buf
        self.shm.close()
        self.shm.unlink()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
