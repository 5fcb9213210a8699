"""Oracle: checkerboard geometry, FD/PML assembly, media and RHS.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates:
* DDCheckerboard::setUp              src/DDCheckerboard.cpp:87-645
* AssemblerHelmholtz2DFD::setN/getDOFPts/getMeshSize
                                     src/AssemblerHelmholtz2DFD.cpp:67-181
* AssemblerHelmholtz2DFD::assembleSystemMatrix / assembleRHS
                                     src/AssemblerHelmholtz2DFD.cpp:347-457
* PMLCubic2D                         src/PMLCubic2D.cpp:35-236
* FDweights / DerivativeDifferenceMatrix1D (Order 2 central stencils)
                                     src/FiniteDifferences.cpp:11-122
* SubdomainStructured index sets     src/SubdomainStructured.cpp:315-397,749-801,1019-1277
* FunctionConstant / FunctionCSV     src/FunctionConstant.cpp:67-70, src/FunctionCSV.cpp:46-236
* RHSSrcPts::eval                    src/RHSSrcPts.cpp:87-112
* Input derived parameters           src/Input.cpp:243,470-623,795-857

Pinning: tests/test_oracle_geometry.py checks the Order-2 Fornberg weights
against the closed forms, interior matrix rows against the textbook
-Lap_h - w^2 m stencil, the PML coefficients against the closed-form alpha and
alpha', the owned-set partition (disjoint, covering the global grid), the
equality of every owned local matrix row with the global operator row, and the
trace/L-trace sets against their geometric definition.
"""
import math

import numpy as np
import scipy.sparse as sp

PI = math.pi  # include/utils.h:12 (M_PI)


# --------------------------------------------------------------------------
# media
# --------------------------------------------------------------------------
class ConstantMedium:
    """Squared slowness 'Constant': always 1.0 (src/Input.cpp:744-751)."""

    def min_real(self):
        return 1.0

    def eval(self, x, y):
        return np.ones(np.broadcast(x, y).shape)


class CSVMedium:
    """FunctionCSV: velocities on a top-row-first grid, m = 1/c^2.

    src/FunctionCSV.cpp:46-236.  Values are parsed as FP32 (``stof``), the
    squared slowness is formed in FP64.  ``setDomain`` crops to the cells
    covering the problem domain; ``eval`` clamps with mH[0] on the low side and
    mH[1] on the high side of *both* axes (:222-227).
    """

    def __init__(self, velocities_f32, file_dom):
        c = np.asarray(velocities_f32, dtype=np.float32)
        self.c = c
        self.file_dom = [list(map(float, file_dom[0])), list(map(float, file_dom[1]))]
        self.nrows, self.ncols = c.shape
        self.mH = [(self.file_dom[0][1] - self.file_dom[0][0]) / self.ncols,
                   (self.file_dom[1][1] - self.file_dom[1][0]) / self.nrows]
        self._set_domain(self.file_dom)

    @classmethod
    def from_file(cls, filename, file_dom):
        rows = []
        with open(filename) as f:
            for line in f:
                line = line.strip()
                if line:
                    rows.append([np.float32(v) for v in line.split(',')])
        return cls(np.array(rows, dtype=np.float32), file_dom)

    def _set_domain(self, dom):
        fd, mH = self.file_dom, self.mH
        d = [[max(dom[0][0], fd[0][0]), min(dom[0][1], fd[0][1])],
             [max(dom[1][0], fd[1][0]), min(dom[1][1], fd[1][1])]]
        start_row = int((fd[1][1] - d[1][1]) / mH[1])
        end_row = int((fd[1][1] - d[1][0]) / mH[1])
        start_col = int((d[0][0] - fd[0][0]) / mH[0])
        end_col = int((d[0][1] - fd[0][0]) / mH[0])
        end_row = min(end_row, self.nrows - 1)
        end_col = min(end_col, self.ncols - 1)
        self.domA = [[fd[0][0] + start_col * mH[0], fd[0][0] + (end_col + 1) * mH[0]],
                     [fd[1][1] - (end_row + 1) * mH[1], fd[1][1] - start_row * mH[1]]]
        cc = self.c[start_row:end_row + 1, start_col:end_col + 1].astype(np.float64)
        self.A = 1.0 / cc ** 2
        self._min = float(self.A.min())

    def set_domain(self, dom):
        self._set_domain(dom)

    def min_real(self):
        return self._min

    def eval(self, x, y):
        x = np.array(x, dtype=np.float64, copy=True)
        y = np.array(y, dtype=np.float64, copy=True)
        mH, dA = self.mH, self.domA
        lo_x = dA[0][0] + 0.5 * mH[0]
        lo_y = dA[1][0] + 0.5 * mH[0]
        hi_x = dA[0][1] - 0.5 * mH[1]
        hi_y = dA[1][1] - 0.5 * mH[1]
        x = np.where(x <= lo_x, lo_x, x)
        y = np.where(y <= lo_y, lo_y, y)
        x = np.where(x >= hi_x, hi_x, x)
        y = np.where(y >= hi_y, hi_y, y)
        col = ((x - dA[0][0]) / mH[0]).astype(np.int64)
        row = ((dA[1][1] - y) / mH[1]).astype(np.int64)
        return self.A[row, col]


# --------------------------------------------------------------------------
# cubic PML (src/PMLCubic2D.cpp:35-236)
# --------------------------------------------------------------------------
def _sigma(v, lo_c, lo_i, hi_i, hi_c, C):
    """sigma and sigma' along one axis for an array of coordinates."""
    s = np.zeros_like(v)
    ds = np.zeros_like(v)
    left = (v >= lo_c) & (v < lo_i)
    right = (v > hi_i) & (v < hi_c)
    inside = (v >= lo_i) & (v <= hi_i)
    if not np.all(left | right | inside):
        raise ValueError("PML evaluation point out of range")  # PMLCubic2D.cpp:142-146
    dl = lo_i - lo_c
    dr = hi_c - hi_i
    s[left] = C * ((lo_i - v[left]) / dl) ** 3.0
    ds[left] = -3.0 * C * (lo_i - v[left]) ** 2.0 / dl ** 3.0
    s[right] = C * ((v[right] - hi_i) / dr) ** 3.0
    ds[right] = 3.0 * C * (v[right] - hi_i) ** 2.0 / dr ** 3.0
    return s, ds


def pml_coefficients(x, y, int_dom, comp_dom, C):
    """Diffusion alpha^2 and convection alpha'*alpha per axis."""
    sx, dsx = _sigma(x, comp_dom[0][0], int_dom[0][0], int_dom[0][1], comp_dom[0][1], C)
    sy, dsy = _sigma(y, comp_dom[1][0], int_dom[1][0], int_dom[1][1], comp_dom[1][1], C)
    a1 = 1.0 / (1.0 + 1j * sx)
    a2 = 1.0 / (1.0 + 1j * sy)
    a1d = -1j * dsx / (1.0 + 1j * sx) ** 2
    a2d = -1j * dsy / (1.0 + 1j * sy) ** 2
    return a1 * a1, a2 * a2, a1d * a1, a2d * a2


def fd_weights(z, xs, m):
    """Fornberg weights (src/FiniteDifferences.cpp:11-54) for derivative m."""
    n = len(xs) - 1
    c = np.zeros((m + 1, n + 1))
    c[0, 0] = 1.0
    c1 = 1.0
    c4 = xs[0] - z
    for i in range(n):
        mn = min(i + 1, m)
        c2 = 1.0
        c5 = c4
        c4 = xs[i + 1] - z
        for j in range(i + 1):
            c3 = xs[i + 1] - xs[j]
            c2 *= c3
            if j == i:
                for k in range(mn - 1, -1, -1):
                    c[k + 1, i + 1] = c1 * ((k + 1) * c[k, i] - c5 * c[k + 1, i]) / c2
                c[0, i + 1] = -c1 * c5 * c[0, i] / c2
            for k in range(mn - 1, -1, -1):
                c[k + 1, j] = (c4 * c[k + 1, j] - (k + 1) * c[k, j]) / c3
            c[0, j] = c4 * c[0, j] / c3
        c1 = c2
    return c[m].copy()


# --------------------------------------------------------------------------
# one subdomain (SubdomainStructured + its AssemblerHelmholtz2DFD)
# --------------------------------------------------------------------------
def local_grid(int_dom, n, pml_width_in):
    """AssemblerHelmholtz2DFD::setN / getMeshSize (src/AssemblerHelmholtz2DFD.cpp:67-144):
    the PML width rounded to the nearest multiple of h = min(h1, h2), the
    computational domain, intervals per axis and the spacings."""
    h1 = (int_dom[0][1] - int_dom[0][0]) / n[0]
    h2 = (int_dom[1][1] - int_dom[1][0]) / n[1]
    h = min(h1, h2)
    ratio = pml_width_in / h
    pml_els = int(pml_width_in / h)
    if ratio - pml_els > 0.5:
        pml_els += 1
    pml = h * pml_els
    cd = [[int_dom[0][0] - pml, int_dom[0][1] + pml],
          [int_dom[1][0] - pml, int_dom[1][1] + pml]]
    mN = [int((cd[0][1] - cd[0][0] + 0.5 * h1) / h1),
          int((cd[1][1] - cd[1][0] + 0.5 * h2) / h2)]
    return dict(pml_els=pml_els, pml_width=pml, comp_dom=cd, mN=mN,
                hx=(cd[0][1] - cd[0][0]) / mN[0], hy=(cd[1][1] - cd[1][0]) / mN[1])


class Cell:
    """One checkerboard subdomain with its local FD/PML grid and index sets."""

    def __init__(self, I, J, int_dom, n, lims1, lims2, pml_width_in, omega, C, medium, index_sets=True):
        self.I, self.J = I, J
        self.int_dom = [list(int_dom[0]), list(int_dom[1])]
        self.omega, self.C, self.medium = omega, C, medium
        g = local_grid(int_dom, n, pml_width_in)
        self.pml_els, self.pml_width, self.comp_dom = g["pml_els"], g["pml_width"], g["comp_dom"]
        self.mN, self.hx, self.hy = g["mN"], g["hx"], g["hy"]
        cd = self.comp_dom
        self.nx, self.ny = self.mN[0] - 1, self.mN[1] - 1     # DOFs per axis
        self.size = self.nx * self.ny
        # getDOFPts: index i*(ny-1)+j, point (x0c+(i+1)hx, y0c+(j+1)hy)
        self.xs = cd[0][0] + (np.arange(self.nx) + 1).astype(np.float64) * self.hx
        self.ys = cd[1][0] + (np.arange(self.ny) + 1).astype(np.float64) * self.hy
        self.mesh_size = math.sqrt(self.hx * self.hy)
        self.A = None
        self.lu = None
        if not index_sets:
            return
        # index sets (assignIndexSets / computeVolumeIndexSets)
        self.lims = [list(lims1), list(lims2)]
        self.col_class = self._classify(self.xs, lims1)
        self.row_class = self._classify(self.ys, lims2)
        self.lr = {s: self._volume_sets(self.col_class, s, axis=0) for s in (False, True)}
        self.ud = {s: self._volume_sets(self.row_class, s, axis=1) for s in (False, True)}
        self.ltr = {}
        for t in ("BottomLeft", "BottomRight", "TopRight", "TopLeft"):
            for slr in (False, True):
                for sud in (False, True):
                    self.ltr[(t, slr, sud)] = self._ltrace_sets(t, slr, sud)

    # -- index sets --------------------------------------------------------
    @staticmethod
    def _classify(coords, lims):
        """domainIndices: first limit the coordinate is below, else 10."""
        cls = np.full(len(coords), 10, dtype=np.int64)
        for k in range(9, -1, -1):
            cls = np.where(coords < lims[k], k, cls)
        return cls

    def _dofs(self, axis_members, axis):
        if axis == 0:
            cols = np.nonzero(axis_members)[0]
            return (cols[:, None] * self.ny + np.arange(self.ny)[None, :]).ravel()
        rows = np.nonzero(axis_members)[0]
        return np.sort((np.arange(self.nx)[:, None] * self.ny + rows[None, :]).ravel())

    def _volume_sets(self, cls, shifted, axis):
        """The 7-set view {below, trace0, trace1, interior, traceN, traceNP,
        above} built from the 11 raw sets (src/SubdomainStructured.cpp:749-801)."""
        groups = ([[0], [1], [2], [3, 4, 5], [6], [7], [8, 9, 10]]
                  if not shifted else
                  [[0, 1, 2], [3], [4], [5, 6, 7], [8], [9], [10]])
        return [self._dofs(np.isin(cls, g), axis) for g in groups]

    def _ltrace_sets(self, t, slr, sud):
        """computeLTraceIndexSets (src/SubdomainStructured.cpp:1119-1277)."""
        s1, s2 = self.lr[slr], self.ud[sud]
        if t == "BottomLeft":
            i0v, i1v, i0h, i1h = s1[1], s1[2], s2[1], s2[2]
            n0v, n0h = s1[0], s2[0]
            n1v, n1h = np.concatenate([s1[0], s1[1]]), np.concatenate([s2[0], s2[1]])
        elif t == "BottomRight":
            i0v, i1v, i0h, i1h = s1[5], s1[4], s2[1], s2[2]
            n0v, n0h = s1[6], s2[0]
            n1v, n1h = np.concatenate([s1[5], s1[6]]), np.concatenate([s2[0], s2[1]])
        elif t == "TopRight":
            i0v, i1v, i0h, i1h = s1[5], s1[4], s2[5], s2[4]
            n0v, n0h = s1[6], s2[6]
            n1v, n1h = np.concatenate([s1[5], s1[6]]), np.concatenate([s2[5], s2[6]])
        else:  # TopLeft
            i0v, i1v, i0h, i1h = s1[1], s1[2], s2[5], s2[4]
            n0v, n0h = s1[0], s2[6]
            n1v, n1h = np.concatenate([s1[0], s1[1]]), np.concatenate([s2[5], s2[6]])
        out = []
        for iv, ih, nv, nh in ((i0v, i0h, n0v, n0h), (i1v, i1h, n1v, n1h)):
            m = np.zeros(self.size, dtype=bool)
            m[iv] = True
            m[ih] = True
            m[nv] = False
            m[nh] = False
            out.append(np.nonzero(m)[0])
        return out

    # -- assembly ------------------------------------------------------------
    def squared_slowness(self):
        X, Y = np.meshgrid(self.xs, self.ys, indexing="ij")
        return self.medium.eval(X.ravel(), Y.ravel())

    def assemble(self):
        """assembleSystemMatrix (src/AssemblerHelmholtz2DFD.cpp:371-457)."""
        self.A = assemble_helmholtz(self.xs, self.ys, self.hx, self.hy, self.int_dom,
                                    self.comp_dom, self.C, self.omega,
                                    self.squared_slowness())
        return self.A

    def factorize(self):
        import scipy.sparse.linalg as spla
        if self.A is None:
            self.assemble()
        self.lu = spla.splu(self.A.tocsc(), permc_spec="COLAMD")

    def solve(self, b):
        return self.lu.solve(np.asarray(b, dtype=np.complex128))

    def dof_points(self):
        X, Y = np.meshgrid(self.xs, self.ys, indexing="ij")
        return X.ravel(), Y.ravel()


def assemble_helmholtz(xs, ys, hx, hy, int_dom, comp_dom, C, omega, m):
    """-(L1 .* D2x(x)I) - (L2 .* I(x)D2y) - (b1 .* D1x(x)I + b2 .* I(x)D1y) - w^2 m.

    Order 2: D2 = [1,-2,1]/h^2 and D1 = [-1/2,0,1/2]/h
    (src/FiniteDifferences.cpp:57-122, block(1,1,n-1,n-1)).  Index i*ny+j.
    """
    nx, ny = len(xs), len(ys)
    w2 = fd_weights(1, [0, 1, 2], 2) / hx ** 2.0
    w1 = fd_weights(1, [0, 1, 2], 1) / hx ** 1.0
    v2 = fd_weights(1, [0, 1, 2], 2) / hy ** 2.0
    v1 = fd_weights(1, [0, 1, 2], 1) / hy ** 1.0
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    L1, L2, B1, B2 = pml_coefficients(X.ravel(), Y.ravel(), int_dom, comp_dom, C)
    N = nx * ny
    idx = np.arange(N)
    i = idx // ny
    j = idx % ny
    diag = -(L1 * w2[1]) - L2 * v2[1] - (B1 * w1[1] + B2 * v1[1]) - omega * omega * m * 1.0
    rows = [idx]
    cols = [idx]
    vals = [diag]
    for di, wa, wb in ((-1, w2[0], w1[0]), (1, w2[2], w1[2])):
        ok = (i + di >= 0) & (i + di < nx)
        rows.append(idx[ok])
        cols.append(idx[ok] + di * ny)
        vals.append(-(L1[ok] * wa) - B1[ok] * wb)
    for dj, va, vb in ((-1, v2[0], v1[0]), (1, v2[2], v1[2])):
        ok = (j + dj >= 0) & (j + dj < ny)
        rows.append(idx[ok])
        cols.append(idx[ok] + dj)
        vals.append(-(L2[ok] * va) - B2[ok] * vb)
    A = sp.csc_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(N, N))
    return A


def point_sources_rhs(xp, yp, sources, eps):
    """RHSSrcPts::eval summed over sources (src/RHSSrcPts.cpp:87-112)."""
    f = np.zeros(len(xp), dtype=np.complex128)
    for (sx, sy) in sources:
        d2 = (sx - xp) ** 2.0 + (sy - yp) ** 2.0
        f += 1.0 / (PI * eps ** 2) * np.exp(-d2 / eps ** 2.0)
    return f


# --------------------------------------------------------------------------
# DDCheckerboard::setUp (src/DDCheckerboard.cpp:87-645)
# --------------------------------------------------------------------------
class Checkerboard:
    """Geometry and local problems of one checkerboard DD (single process).

    ``domain`` = [[ax,bx],[ay,by]], ``ncells`` = (NCellsX, NCellsY) sizing
    rows and columns, ``nels`` = (NX, NY).  Cells are stored row-major
    (I*nC+J), which is mCells order for one MPI rank.
    """

    def __init__(self, domain, ncells, nels, omega, min_wl_in_pml, medium, order=2,
                 assemble=True, general=False, cells=True):
        """cells=False keeps only what the global grid needs: the global FD/PML
        operator, the DD mesh size (from each cell's local grid spacing) and
        the global right-hand side; no cell index sets or local matrices."""
        if order != 2:
            raise ValueError("only Order 2 is supported")
        tt = int(math.ceil(order / 2.0))
        if nels[0] <= 4 * tt or nels[1] <= 4 * tt:
            raise ValueError("Not enough elements in cell!")  # :108-112
        self.domain = [list(map(float, domain[0])), list(map(float, domain[1]))]
        self.nR, self.nC = int(ncells[0]), int(ncells[1])
        self.nels = (int(nels[0]), int(nels[1]))
        self.omega = float(omega)
        self.medium = medium
        # Input::readAssemblerHelmholtz2DFD (src/Input.cpp:834-847)
        self.C = math.log(self.omega)
        dom = self.domain
        wl1 = (dom[0][1] - dom[0][0]) * 2 * PI / self.omega
        wl2 = (dom[1][1] - dom[1][0]) * 2 * PI / self.omega
        wl = min(wl1, wl2)
        wl /= math.sqrt(medium.min_real())
        self.pml_width_in = min_wl_in_pml * wl
        if general:
            # per-axis spacing: x from the columns and NY, y from the rows and NX
            # (coincides with the reference's rule when nR(NX+1) == nC(NY+1))
            n1 = self.nC * self.nels[1] + self.nC - 1
            n2 = self.nR * self.nels[0] + self.nR - 1
        else:
            n1 = self.nR * self.nels[0] + self.nR - 1
            n2 = self.nC * self.nels[1] + self.nC - 1
        self.n_el = (n1, n2)
        h1 = (dom[0][1] - dom[0][0]) / n1
        h2 = (dom[1][1] - dom[1][0]) / n2
        self.h = (h1, h2)
        NX, NY = self.nels
        self.cells = []
        self._cell_mesh = []
        for I in range(self.nR):
            for J in range(self.nC):
                lims1 = [0.0] * 10
                lims2 = [0.0] * 10
                a = dom[0][0] + (NY + 1) * J * h1
                if J > 0:
                    lims1[0] = a - tt * h1 - 0.5 * h1
                    lims1[1] = a - 0.5 * h1
                    lims1[2] = a + tt * h1 - 0.5 * h1
                    lims1[3] = a + 2 * tt * h1 - 0.5 * h1
                    lims1[4] = a + 3 * tt * h1 - 0.5 * h1
                    a -= 2 * tt * h1
                    if a < dom[0][0]:
                        a = dom[0][0]
                b = dom[0][0] + (NY + 1) * J * h1 + NY * h1
                if J < self.nC - 1:
                    lims1[5] = b - tt * h1 + 0.5 * h1
                    lims1[6] = b + 0.5 * h1
                    lims1[7] = b + tt * h1 + 0.5 * h1
                    lims1[8] = b + 2 * tt * h1 + 0.5 * h1
                    lims1[9] = b + 3 * tt * h1 + 0.5 * h1
                    b += 4 * tt * h1
                    if b > dom[0][1]:
                        b = dom[0][1]
                int_dom = [[a, b], [0.0, 0.0]]
                a = dom[1][0] + (NX + 1) * I * h2
                if I > 0:
                    lims2[0] = a - tt * h2 - 0.5 * h2
                    lims2[1] = a - 0.5 * h2
                    lims2[2] = a + tt * h2 - 0.5 * h2
                    lims2[3] = a + 2 * tt * h2 - 0.5 * h2
                    lims2[4] = a + 3 * tt * h2 - 0.5 * h2
                    a -= 2 * tt * h2
                    if a < dom[1][0]:
                        a = dom[1][0]
                b = dom[1][0] + (NX + 1) * I * h2 + NX * h2
                if I < self.nR - 1:
                    lims2[5] = b - tt * h2 + 0.5 * h2
                    lims2[6] = b + 0.5 * h2
                    lims2[7] = b + tt * h2 + 0.5 * h2
                    lims2[8] = b + 2 * tt * h2 + 0.5 * h2
                    lims2[9] = b + 3 * tt * h2 + 0.5 * h2
                    b += 4 * tt * h2
                    if b > dom[1][1]:
                        b = dom[1][1]
                int_dom[1] = [a, b]
                nloc = [int((int_dom[0][1] - int_dom[0][0] + 0.5 * h1) / h1),
                        int((int_dom[1][1] - int_dom[1][0] + 0.5 * h2) / h2)]
                if not cells:
                    g = local_grid(int_dom, nloc, self.pml_width_in)
                    self._cell_mesh.append(math.sqrt(g["hx"] * g["hy"]))
                    continue
                # provisional cell to get the DOF points for the boundary limits
                probe = Cell(I, J, int_dom, nloc, [0.0] * 10, [0.0] * 10,
                             self.pml_width_in, self.omega, self.C, medium)
                # quirky min/max scan starting at 0.0 with else-if (:347-361)
                minx = maxx = miny = maxy = 0.0
                X, Y = probe.dof_points()
                for xv, yv in zip(X, Y):
                    if xv < minx:
                        minx = xv
                    elif xv > maxx:
                        maxx = xv
                    if yv < miny:
                        miny = yv
                    elif yv > maxy:
                        maxy = yv
                if J == 0:
                    lims1[0:5] = [minx - h1] * 5
                if J == self.nC - 1:
                    lims1[5:10] = [maxx + h1] * 5
                if I == 0:
                    lims2[0:5] = [miny - h2] * 5
                if I == self.nR - 1:
                    lims2[5:10] = [maxy + h2] * 5
                cell = Cell(I, J, int_dom, nloc, lims1, lims2, self.pml_width_in,
                            self.omega, self.C, medium)
                self.cells.append(cell)
        for c in self.cells:
            self._owned_sets(c)
        if assemble:
            for c in self.cells:
                c.assemble()
        self._global_grid()

    def cell(self, I, J):
        return self.cells[I * self.nC + J]

    # mValsToKeep* (src/DDCheckerboard.cpp:434-634)
    @staticmethod
    def _owned_sets(c):
        def keep(lr, ud):
            m = np.zeros(c.size, dtype=bool)
            for k in (2, 3, 4):
                m[lr[k]] = True
            for k in (0, 1, 5, 6):
                m[ud[k]] = False
            return np.nonzero(m)[0]
        c.keep = {
            (False, False): keep(c.lr[False], c.ud[False]),
            (True, True): keep(c.lr[True], c.ud[True]),
            (True, False): keep(c.lr[True], c.ud[False]),
            (False, True): keep(c.lr[False], c.ud[True]),
        }

    def mesh_size(self):
        """DDCheckerboard::getMeshSize (src/DDCheckerboard.cpp:3536-3561)."""
        if not self.cells:
            return min(self._cell_mesh)
        return min(c.mesh_size for c in self.cells)

    # ------------------------------------------------------------------
    # global grid: the operator of a 1x1 DD over the whole domain
    # ------------------------------------------------------------------
    def _global_grid(self):
        n1, n2 = self.n_el
        h1, h2 = self.h
        dom = self.domain
        h = min(h1, h2)
        ratio = self.pml_width_in / h
        pml_els = int(self.pml_width_in / h)
        if ratio - pml_els > 0.5:
            pml_els += 1
        self.pml_els = pml_els
        # global DOF grid = the local grid of a 1x1 DD (src/AssemblerHelmholtz2DFD.cpp:166-177)
        g = Cell(0, 0, dom, [n1, n2], [0.0] * 10, [0.0] * 10, self.pml_width_in,
                 self.omega, self.C, self.medium, index_sets=False)
        self.Gx, self.Gy = g.nx, g.ny
        self.N = self.Gx * self.Gy
        self.gcell = g
        self.gxs, self.gys = g.xs, g.ys
        # local -> global offsets (grid-aligned)
        for c in self.cells:
            ox = int(round((c.xs[0] - g.xs[0]) / h1))
            oy = int(round((c.ys[0] - g.ys[0]) / h2))
            c.ox, c.oy = ox, oy
            ii = np.arange(c.nx)[:, None] + ox
            jj = np.arange(c.ny)[None, :] + oy
            c.gidx = (ii * self.Gy + jj).ravel()   # may be out of range in local PML

    def global_operator(self):
        if self.gcell.A is None:
            self.gcell.assemble()
        return self.gcell.A

    def global_owned_index(self, shifted=(False, False)):
        """For each cell, the global DOF index of each owned entry."""
        return [c.gidx[c.keep[shifted]] for c in self.cells]

    def to_global(self, owned_vecs):
        y = np.zeros(self.N, dtype=np.complex128)
        for c, v in zip(self.cells, owned_vecs):
            y[c.gidx[c.keep[(False, False)]]] = v
        return y

    def from_global(self, y):
        return [np.asarray(y[c.gidx[c.keep[(False, False)]]], dtype=np.complex128)
                for c in self.cells]

    # DDCheckerboard::assembleRHS (src/DDCheckerboard.cpp:668-684)
    def assemble_rhs(self, sources):
        eps = self.mesh_size()
        out = []
        for c in self.cells:
            X, Y = c.dof_points()
            f = point_sources_rhs(X, Y, sources, eps)
            out.append(f[c.keep[(False, False)]])
        return out

    def global_rhs(self, sources):
        if not self.cells:
            # cell DOF points are global DOF points (grid-aligned cells)
            X, Y = self.gcell.dof_points()
            return point_sources_rhs(X, Y, sources, self.mesh_size())
        return self.to_global(self.assemble_rhs(sources))

    def factorize(self):
        for c in self.cells:
            c.factorize()
