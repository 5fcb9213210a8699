"""Oracle: the L-sweeps preconditioner and the global operator.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates, on one process, with every cell held in a Python list in mCells
order (row-major, I*nC+J):

* DDCheckerboard::applyPrecond          src/DDCheckerboard.cpp:687-739
* getSeparatedRHS                       :3564-3674
* reduce / expand                       :744-809
* applySweep                            :812-880
* addLocSol (stage 1)                   :883-973
* addLeftSweep / addRightSweep          :976-1164 / :1168-1350
* addDownSweep / addUpSweep             :1353-1540 / :1544-1724
* addTL2BRSweep / addTR2BLSweep         :1727-1946 / :1949-2170
* addBR2TLSweep / addBL2TRSweep         :2173-2390 / :2393-2608
* convert{NonShiftedToShifted,ShiftedToNonShifted}{LR,UD}
                                        :2995-3515
* SubdomainStructured::getEquivalentSources / extractLTraces
                                        src/SubdomainStructured.cpp:1304-1539
* DDCheckerboard::applySystemMatrix     :2611-2635 (owned rows of the local
  matrices are rows of the global operator, so it is the global matvec)

The cell processing order inside each sweep follows getSortedIndices on the
reference's diagonal index (src/utils.cpp:242-260).  MPI messages between
rows become list accesses.

Pinning: tests/test_oracle_sweeps.py checks that on a 1x1 board the
preconditioner is the exact inverse (all trace sets are empty), that it is
linear, that the shifted/unshifted conversions are exact re-layouts of one
global grid function, and that GMRES preconditioned by it converges to the
global direct solution.
"""
import numpy as np

FF, TF, FT, TT = (False, False), (True, False), (False, True), (True, True)


def _order(cb, flip_i, flip_j):
    """getSortedIndices on the (optionally mirrored) diagonal index."""
    return [idx for batch in _diagonals(cb, flip_i, flip_j) for idx in batch]


def _diagonals(cb, flip_i, flip_j):
    """The processing order of getSortedIndices cut into its anti-diagonals:
    a sweep reads only the traces of cells on earlier diagonals, so the
    local solves of one diagonal are independent (the reference's ranks run
    them at the same time)."""
    diag = []
    for idx, c in enumerate(cb.cells):
        I = cb.nR - 1 - c.I if flip_i else c.I
        J = cb.nC - 1 - c.J if flip_j else c.J
        diag.append((float(I + J), idx))
    diag.sort()
    out = []
    for d, idx in diag:
        if out and out[-1][0] == d:
            out[-1][1].append(idx)
        else:
            out.append((d, [idx]))
    return [b for _, b in out]


class Preconditioner:
    """L-sweeps M^{-1} over one factorized Checkerboard."""

    def __init__(self, cb, solver=None):
        """solver (optional): solve_many([(cell index, rhs), ...]) -> list of
        solutions, e.g. a process pool holding the factors; by default each
        cell's own factors, one after the other."""
        self.cb = cb
        self.nsolves = 0
        self.solver = solver

    # -- helpers -------------------------------------------------------------
    def expand(self, x, shift=FF):
        out = []
        for c, v in zip(self.cb.cells, x):
            e = np.zeros(c.size, dtype=np.complex128)
            e[c.keep[shift]] = v
            out.append(e)
        return out

    def reduce(self, x, shift=FF):
        return [np.array(v[c.keep[shift]]) for c, v in zip(self.cb.cells, x)]

    def _solve_many(self, reqs):
        self.nsolves += len(reqs)
        if self.solver is not None:
            return self.solver.solve_many(reqs)
        return [self.cb.cells[i].solve(r) for i, r in reqs]

    def _run(self, batches, body):
        """Runs body(idx) for every cell, a diagonal at a time.  body is a
        generator that yields the right-hand side of its local solve (or
        returns without one) and receives the solution; the solves of one
        diagonal go to _solve_many together."""
        for batch in batches:
            gens, reqs = [], []
            for idx in batch:
                g = body(idx)
                try:
                    rhs = next(g)
                except StopIteration:
                    continue
                gens.append(g)
                reqs.append((idx, rhs))
            for g, sol in zip(gens, self._solve_many(reqs) if reqs else []):
                try:
                    g.send(sol)
                except StopIteration:
                    pass

    def _idx(self, I, J):
        return I * self.cb.nC + J

    def separated_rhs(self, x, slr, sud):
        out = self.expand(x, FF)
        for c, e in zip(self.cb.cells, out):
            lr, ud = c.lr[False], c.ud[False]
            if not slr and not sud:
                zl, zu = (1, 2, 4, 5), (1, 2, 4, 5)
            elif not slr:
                zl, zu = (1, 2, 4, 5), (0, 3, 6)
            elif not sud:
                zl, zu = (0, 3, 6), (1, 2, 4, 5)
            else:
                zl, zu = (0, 3, 6), (0, 3, 6)
            for k in zl:
                e[lr[k]] = 0.0
            for k in zu:
                e[ud[k]] = 0.0
        return out

    # -- layout conversions (exact re-layouts of one grid function) ----------
    def nonshifted_to_shifted_lr(self, x, sud):
        cb = self.cb
        xe = self.expand(x, (False, sud))
        for idx in _order(cb, True, True):
            c = cb.cells[idx]
            if c.J < cb.nC - 1:
                r = cb.cells[self._idx(c.I, c.J + 1)]
                xe[idx][c.lr[False][5]] = xe[self._idx(c.I, c.J + 1)][r.lr[False][2]]
                xe[idx][c.lr[True][4]] = xe[self._idx(c.I, c.J + 1)][r.lr[True][1]]
        return self.reduce(xe, (True, sud))

    def nonshifted_to_shifted_ud(self, x, slr):
        cb = self.cb
        xe = self.expand(x, (slr, False))
        for idx in _order(cb, True, True):
            c = cb.cells[idx]
            if c.I < cb.nR - 1:
                t = cb.cells[self._idx(c.I + 1, c.J)]
                xe[idx][c.ud[False][5]] = xe[self._idx(c.I + 1, c.J)][t.ud[False][2]]
                xe[idx][c.ud[True][4]] = xe[self._idx(c.I + 1, c.J)][t.ud[True][1]]
        return self.reduce(xe, (slr, True))

    def shifted_to_nonshifted_lr(self, x, sud):
        cb = self.cb
        xe = self.expand(x, (True, sud))
        for idx in _order(cb, False, False):
            c = cb.cells[idx]
            if c.J > 0:
                left = cb.cells[self._idx(c.I, c.J - 1)]
                li = self._idx(c.I, c.J - 1)
                xe[idx][c.lr[True][1]] = xe[li][left.lr[True][4]]
                xe[idx][c.lr[False][2]] = xe[li][left.lr[False][5]]
        return self.reduce(xe, (False, sud))

    def shifted_to_nonshifted_ud(self, x, slr):
        cb = self.cb
        xe = self.expand(x, (slr, True))
        for idx in _order(cb, False, False):
            c = cb.cells[idx]
            if c.I > 0:
                b = cb.cells[self._idx(c.I - 1, c.J)]
                bi = self._idx(c.I - 1, c.J)
                xe[idx][c.ud[True][1]] = xe[bi][b.ud[True][4]]
                xe[idx][c.ud[False][2]] = xe[bi][b.ud[False][5]]
        return self.reduce(xe, (slr, False))

    # -- equivalent sources --------------------------------------------------
    @staticmethod
    def equiv_sources(c, kind, slr, sud, vals0, vals1):
        lr, ud = c.lr[slr], c.ud[sud]
        if kind == "Bottom":
            r0, r1 = ud[1], ud[2]
        elif kind == "Right":
            r0, r1 = lr[5], lr[4]
        elif kind == "Top":
            r0, r1 = ud[5], ud[4]
        elif kind == "Left":
            r0, r1 = lr[1], lr[2]
        else:
            r0, r1 = c.ltr[(kind, slr, sud)]
        # the coupling blocks A01 = A[trace0, trace1], A10 = A[trace1, trace0]
        # are built once per cell, kind and shift, as the reference builds
        # them at set-up (src/SubdomainStructured.cpp:549-654)
        cache = c.__dict__.setdefault("_coupling", {})
        key = (kind, slr, sud)
        if key not in cache:
            A = c.A.tocsr()
            cache[key] = (A[r0][:, r1].tocsr(), A[r1][:, r0].tocsr())
        A01, A10 = cache[key]
        return A01 @ vals1, -(A10 @ vals0)

    @staticmethod
    def extract_ltraces(c, kind, slr, sud, lr0v, lr1v, ud0v, ud1v):
        lr, ud = c.lr[slr], c.ud[sud]
        if kind == "BottomLeft":
            iL0, iL1, iU0, iU1 = lr[1], lr[2], ud[1], ud[2]
        elif kind == "BottomRight":
            iL0, iL1, iU0, iU1 = lr[5], lr[4], ud[1], ud[2]
        elif kind == "TopRight":
            iL0, iL1, iU0, iU1 = lr[5], lr[4], ud[5], ud[4]
        else:
            iL0, iL1, iU0, iU1 = lr[1], lr[2], ud[5], ud[4]
        t = np.zeros(c.size, dtype=np.complex128)
        t[iL1] = lr1v
        t[iU1] = ud1v
        t[iL0] = lr0v
        t[iU0] = ud0v
        s0, s1 = c.ltr[(kind, slr, sud)]
        return t[s0], t[s1]

    # -- the sweep -------------------------------------------------------------
    def apply_sweep(self, x, slr, sud):
        cb = self.cb
        shift = (slr, sud)
        xt = self.expand(x, shift)
        res = [np.zeros(c.size, dtype=np.complex128) for c in cb.cells]
        LRloc, UDloc = self.stage1(xt, res, slr, sud)
        UD_right = self.stage2_lr(res, LRloc, slr, sud, rightward=True)
        UD_left = self.stage2_lr(res, LRloc, slr, sud, rightward=False)
        LR_up = self.stage2_ud(res, UDloc, slr, sud, upward=True)
        LR_down = self.stage2_ud(res, UDloc, slr, sud, upward=False)
        self.stage3(res, LR_up, UD_right, slr, sud, "BottomLeft")
        self.stage3(res, LR_down, UD_left, slr, sud, "TopRight")
        self.stage3(res, LR_up, UD_left, slr, sud, "BottomRight")
        self.stage3(res, LR_down, UD_right, slr, sud, "TopLeft")
        return self.reduce(res, shift)

    def stage1(self, x, res, slr, sud):
        n = len(self.cb.cells)
        LR = [[None] * n for _ in range(4)]
        UD = [[None] * n for _ in range(4)]

        def body(idx):
            c = self.cb.cells[idx]
            s = yield x[idx]
            res[idx] += s
            lr, ud = c.lr[slr], c.ud[sud]
            for k, si in enumerate((1, 2, 4, 5)):
                UD[k][idx] = s[ud[si]].copy()
                LR[k][idx] = s[lr[si]].copy()

        self._run([list(range(n))], body)
        return LR, UD

    def stage2_lr(self, res, LRin, slr, sud, rightward):
        """addRightSweep (rightward) or addLeftSweep."""
        cb = self.cb
        n = len(cb.cells)
        LR = [[None if v is None else v.copy() for v in LRin[k]] for k in range(4)]
        UD = [[None] * n for _ in range(4)]
        batches = _diagonals(cb, False, False) if rightward else _diagonals(cb, True, True)

        def body(idx):
            c = cb.cells[idx]
            lr, ud = c.lr[slr], c.ud[sud]
            for k, si in enumerate((1, 2, 4, 5)):
                UD[k][idx] = np.zeros(len(ud[si]), dtype=np.complex128)
            if rightward and c.J > 0:
                li = self._idx(c.I, c.J - 1)
                LR[0][idx] = LR[2][li].copy()
                LR[1][idx] = LR[3][li].copy()
                e0, e1 = self.equiv_sources(c, "Left", slr, sud, LR[0][idx], LR[1][idx])
                rhs = np.zeros(c.size, dtype=np.complex128)
                rhs[lr[1]] = e0
                rhs[lr[2]] = e1
            elif (not rightward) and c.J < cb.nC - 1:
                ri = self._idx(c.I, c.J + 1)
                LR[2][idx] = LR[0][ri].copy()
                LR[3][idx] = LR[1][ri].copy()
                e0, e1 = self.equiv_sources(c, "Right", slr, sud, LR[3][idx], LR[2][idx])
                rhs = np.zeros(c.size, dtype=np.complex128)
                rhs[lr[4]] = e1
                rhs[lr[5]] = e0
            else:
                return
            s = yield rhs
            res[idx] += s
            for k, si in enumerate((1, 2, 4, 5)):
                UD[k][idx] += s[ud[si]]
            if rightward:
                LR[2][idx] += s[lr[4]]
                LR[3][idx] += s[lr[5]]
            else:
                LR[0][idx] += s[lr[1]]
                LR[1][idx] += s[lr[2]]

        self._run(batches, body)
        return UD

    def stage2_ud(self, res, UDin, slr, sud, upward):
        """addUpSweep (upward) or addDownSweep."""
        cb = self.cb
        n = len(cb.cells)
        UD = [[None if v is None else v.copy() for v in UDin[k]] for k in range(4)]
        LR = [[None] * n for _ in range(4)]
        batches = _diagonals(cb, False, False) if upward else _diagonals(cb, True, True)

        def body(idx):
            c = cb.cells[idx]
            lr, ud = c.lr[slr], c.ud[sud]
            for k, si in enumerate((1, 2, 4, 5)):
                LR[k][idx] = np.zeros(len(lr[si]), dtype=np.complex128)
            if upward and c.I > 0:
                bi = self._idx(c.I - 1, c.J)
                UD[0][idx] = UD[2][bi].copy()
                UD[1][idx] = UD[3][bi].copy()
                e0, e1 = self.equiv_sources(c, "Bottom", slr, sud, UD[0][idx], UD[1][idx])
                rhs = np.zeros(c.size, dtype=np.complex128)
                rhs[ud[1]] = e0
                rhs[ud[2]] = e1
            elif (not upward) and c.I < cb.nR - 1:
                ti = self._idx(c.I + 1, c.J)
                UD[2][idx] = UD[0][ti].copy()
                UD[3][idx] = UD[1][ti].copy()
                e0, e1 = self.equiv_sources(c, "Top", slr, sud, UD[3][idx], UD[2][idx])
                rhs = np.zeros(c.size, dtype=np.complex128)
                rhs[ud[4]] = e1
                rhs[ud[5]] = e0
            else:
                return
            s = yield rhs
            res[idx] += s
            for k, si in enumerate((1, 2, 4, 5)):
                LR[k][idx] += s[lr[si]]
            if upward:
                UD[2][idx] += s[ud[4]]
                UD[3][idx] += s[ud[5]]
            else:
                UD[0][idx] += s[ud[1]]
                UD[1][idx] += s[ud[2]]

        self._run(batches, body)
        return LR

    def stage3(self, res, LRin, UDin, slr, sud, kind):
        """The four diagonal L-sweeps; the trace sets are taken by value."""
        cb = self.cb
        LR = [[v.copy() for v in LRin[k]] for k in range(4)]
        UD = [[v.copy() for v in UDin[k]] for k in range(4)]
        from_left = kind in ("BottomLeft", "TopLeft")
        from_bottom = kind in ("BottomLeft", "BottomRight")
        # diagonal index mirrors (BL2TR: none; TR2BL: both; BR2TL: J; TL2BR: I)
        batches = _diagonals(cb, not from_bottom, not from_left)

        def body(idx):
            c = cb.cells[idx]
            I, J = c.I, c.J
            okI = I > 0 if from_bottom else I < cb.nR - 1
            okJ = J > 0 if from_left else J < cb.nC - 1
            if not (okI and okJ):
                return
            lr, ud = c.lr[slr], c.ud[sud]
            if from_left:
                li = self._idx(I, J - 1)
                LR[0][idx] = LR[2][li].copy()
                LR[1][idx] = LR[3][li].copy()
                vlr0, vlr1 = LR[0][idx], LR[1][idx]
            else:
                ri = self._idx(I, J + 1)
                LR[2][idx] = LR[0][ri].copy()
                LR[3][idx] = LR[1][ri].copy()
                vlr0, vlr1 = LR[3][idx], LR[2][idx]
            if from_bottom:
                bi = self._idx(I - 1, J)
                UD[0][idx] = UD[2][bi].copy()
                UD[1][idx] = UD[3][bi].copy()
                vud0, vud1 = UD[0][idx], UD[1][idx]
            else:
                ti = self._idx(I + 1, J)
                UD[2][idx] = UD[0][ti].copy()
                UD[3][idx] = UD[1][ti].copy()
                vud0, vud1 = UD[3][idx], UD[2][idx]
            l0, l1 = self.extract_ltraces(c, kind, slr, sud, vlr0, vlr1, vud0, vud1)
            e0, e1 = self.equiv_sources(c, kind, slr, sud, l0, l1)
            s0, s1 = c.ltr[(kind, slr, sud)]
            rhs = np.zeros(c.size, dtype=np.complex128)
            rhs[s0] = e0
            rhs[s1] = e1
            s = yield rhs
            if kind != "BottomLeft":
                res[idx] += s
            # overwrite the incoming outer traces into the solution
            s = s.copy()
            if from_left:
                s[lr[1]] = LR[0][idx]
            else:
                s[lr[5]] = LR[3][idx]
            if from_bottom:
                s[ud[1]] = UD[0][idx]
            else:
                s[ud[5]] = UD[3][idx]
            if kind == "BottomLeft":
                res[idx] += s
            if from_left:
                LR[2][idx] += s[lr[4]]
                LR[3][idx] += s[lr[5]]
            else:
                LR[0][idx] += s[lr[1]]
                LR[1][idx] += s[lr[2]]
            if from_bottom:
                UD[2][idx] += s[ud[4]]
                UD[3][idx] += s[ud[5]]
            else:
                UD[0][idx] += s[ud[1]]
                UD[1][idx] += s[ud[2]]

        self._run(batches, body)

    # -- M^{-1} ----------------------------------------------------------------
    def apply(self, x):
        """applyPrecond on owned-DOF vectors (list per cell, mCells order)."""
        rhs1 = self.separated_rhs(x, False, False)
        rhs2 = self.separated_rhs(x, True, False)
        rhs3 = self.separated_rhs(x, False, True)
        rhs4 = self.separated_rhs(x, True, True)
        r = self.reduce(rhs1, FF)
        res1 = self.apply_sweep(r, False, False)
        r = self.reduce(rhs2, FF)
        r = self.nonshifted_to_shifted_lr(r, False)
        res2 = self.apply_sweep(r, True, False)
        res2 = self.shifted_to_nonshifted_lr(res2, False)
        r = self.reduce(rhs3, FF)
        r = self.nonshifted_to_shifted_ud(r, False)
        res3 = self.apply_sweep(r, False, True)
        res3 = self.shifted_to_nonshifted_ud(res3, False)
        r = self.reduce(rhs4, FF)
        r = self.nonshifted_to_shifted_lr(r, False)
        r = self.nonshifted_to_shifted_ud(r, True)
        res4 = self.apply_sweep(r, True, True)
        res4 = self.shifted_to_nonshifted_ud(res4, True)
        res4 = self.shifted_to_nonshifted_lr(res4, False)
        return [a + (b + c + d) for a, b, c, d in zip(res1, res2, res3, res4)]

    def apply_global(self, y):
        """M^{-1} on a global grid vector (the GPU path's representation)."""
        return self.cb.to_global(self.apply(self.cb.from_global(y)))


class Operator:
    """applySystemMatrix as the global FD operator on owned-DOF vectors."""

    def __init__(self, cb):
        self.cb = cb
        self.A = cb.global_operator()

    def apply(self, x):
        return self.cb.from_global(self.A @ self.cb.to_global(x))

    def apply_global(self, y):
        return self.A @ y
