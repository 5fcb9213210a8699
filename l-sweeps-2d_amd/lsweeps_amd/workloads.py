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


def layered_velocity(nx, ny, Lx, Ly, seed, n_layers=10, vmin=1.5, vmax=4.5, smooth_cells=5.0,
                     field_std=0.05):
    """Seeded layered model on nx x ny cells over [0,Lx]x[0,Ly] (row 0 = top).

    n_layers-1 interface depths drawn uniform in (0, Ly) and sorted; each
    interface is perturbed by 3 sinusoids of random amplitude (at most 4% of
    its depth), frequency and phase.  Layer velocities rise linearly from vmin
    to vmax plus U(-0.3, 0.3); a Gaussian-smoothed (sigma = smooth_cells)
    random field scales the result by (1 + field) with field std field_std.
    """
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(seed)
    depths = np.sort(rng.uniform(0.0, Ly, n_layers - 1))
    x = (np.arange(nx) + 0.5) * (Lx / nx)
    z = (np.arange(ny) + 0.5) * (Ly / ny)           # depth below the top, row order
    ifaces = []
    for d in depths:
        zk = np.full(nx, d)
        for _ in range(3):
            amp = rng.uniform(0.0, 0.04 * d)
            freq = rng.uniform(0.5, 3.0)
            ph = rng.uniform(0.0, 2 * math.pi)
            zk = zk + amp * np.sin(2 * math.pi * freq * x / Lx + ph)
        ifaces.append(zk)
    ifaces = np.array(ifaces)                        # (n_layers-1, nx)
    layer = (z[:, None, None] > ifaces.T[None, :, :]).sum(axis=2)  # (ny, nx)
    v = vmin + (vmax - vmin) * np.arange(n_layers) / (n_layers - 1) + rng.uniform(-0.3, 0.3, n_layers)
    c = v[layer]
    f = gaussian_filter(rng.standard_normal((ny, nx)), smooth_cells, mode="nearest")
    f *= field_std / max(f.std(), 1e-30)
    c = c * (1.0 + f)
    return c.astype(np.float32)


def config_layered_3x1(seed=0, nsrc=128, ppw=8.0, cells=(4, 12), cell_nodes=200):
    """The layered 3x1 workload: 12 x 4 cells of 200 x 200 nodes.

    The reference tiling needs n = cells*(nodes per cell) - 1 intervals per
    axis, so 2401 x 801 nodes at h = 1/800 cannot be tiled by 12 x 4 equal
    cells; the closest tiling is 2400 x 800 nodes at h = 1/799 on
    [0, 2399/799] x [0, 1] (general geometry mode).  omega gives `ppw` points
    per wavelength at the minimum velocity, 20 PML cells, `nsrc` sources at
    depth 10 cells.
    """
    nR, nC = cells
    nyn, nxn = nR * cell_nodes, nC * cell_nodes
    h = 1.0 / (nyn - 1)
    Lx, Ly = (nxn - 1) * h, 1.0
    vel = layered_velocity(nxn, nyn, Lx + h, Ly + h, seed)   # cells centred on the nodes
    vdom = ((-h / 2, Lx + h / 2), (-h / 2, Ly + h / 2))
    cmin, cmax = float(vel.min()), float(vel.max())
    omega = 2 * math.pi * cmin / (ppw * h)
    wl = 20 * h * omega / (2 * math.pi * cmax)
    xs = (np.arange(nsrc) + 0.5) * Lx / nsrc
    srcs = [(float(x), Ly - 10 * h) for x in xs]
    return dict(name=f"layered{nxn}x{nyn}_s{seed}", domain=((0.0, Lx), (0.0, Ly)), ncells=(nR, nC),
                nels=(cell_nodes - 1, cell_nodes - 1), omega=omega, min_wl_in_pml=wl, velocity=vel,
                vel_domain=vdom, sources=srcs, tol=1e-6, max_iter=60, grid=(nxn, nyn), general=True,
                ppw=ppw, seed=seed)


def _rect_problem(nxn, nyn, cells, vel_fn, ppw, pml_cells, sources_fn, name, tol=1e-6, max_iter=200):
    """A general-geometry problem on nxn x nyn nodes with h = 1/(nyn-1):
    domain [0, (nxn-1)h] x [0, 1], cells = (rows, columns) equal cells."""
    nR, nC = cells
    assert nxn % nC == 0 and nyn % nR == 0
    h = 1.0 / (nyn - 1)
    Lx = (nxn - 1) * h
    vel = vel_fn(nxn, nyn, Lx + h, 1.0 + h)
    cmin, cmax = float(vel.min()), float(vel.max())
    omega = 2 * math.pi * cmin / (ppw * h)
    wl = pml_cells * h * omega / (2 * math.pi * cmax)
    return dict(name=name, domain=((0.0, Lx), (0.0, 1.0)), ncells=(nR, nC),
                nels=(nyn // nR - 1, nxn // nC - 1), omega=omega, min_wl_in_pml=wl, velocity=vel,
                vel_domain=((-h / 2, Lx + h / 2), (-h / 2, 1.0 + h / 2)), sources=sources_fn(Lx, h),
                tol=tol, max_iter=max_iter, grid=(nxn, nyn), general=True, ppw=ppw)


def lens_velocity_rect(nx, ny, Lx, Ly, cx, cy, amp, width=0.02):
    """Gaussian lens c = 1 - amp exp(-r^2/width) on nx x ny cells over
    [0,Lx]x[0,Ly] (row 0 = top), centre (cx, cy) in domain units."""
    xs = (np.arange(nx) + 0.5) * (Lx / nx)
    ys = Ly - (np.arange(ny) + 0.5) * (Ly / ny)
    X, Y = np.meshgrid(xs, ys)
    return (1.0 - amp * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / width)).astype(np.float32)


def config_pack48(seed=0, count=48):
    """The uneven batch: `count` independent problems drawn from one seed.

    Each problem: the long grid side drawn from {256, 384, ..., 1536}, the
    short side from the same set within aspect ratio 1..3, orientation
    random; cells of about 128-200 nodes per side (2..10 per axis); either a
    Gaussian lens (random centre and amplitude) or a layered model (the
    layered generator with a per-problem seed); 8-12 points per wavelength;
    20 PML cells; 1-16 point sources at random interior positions.
    """
    rng = np.random.default_rng(seed)
    sides = np.arange(256, 1537, 128)
    out = []
    for k in range(count):
        L = int(rng.choice(sides))
        shorts = sides[(sides >= L / 3.0) & (sides <= L)]
        S = int(rng.choice(shorts))
        wide = bool(rng.integers(0, 2))
        nxn, nyn = (L, S) if wide else (S, L)
        cpx = int(np.clip(round(nxn / rng.uniform(128, 200)), 2, 10))
        cpy = int(np.clip(round(nyn / rng.uniform(128, 200)), 2, 10))
        nxn -= nxn % cpx
        nyn -= nyn % cpy
        ppw = float(rng.uniform(8.0, 12.0))
        ns = int(rng.integers(1, 17))
        lens = bool(rng.integers(0, 2))
        pseed = int(rng.integers(0, 2**31 - 1))
        cx, cy, amp = float(rng.uniform(0.3, 0.7)), float(rng.uniform(0.3, 0.7)), float(rng.uniform(0.2, 0.5))
        sx, sy = rng.uniform(0.1, 0.9, ns), rng.uniform(0.1, 0.9, ns)
        if lens:
            vel_fn = (lambda nx, ny, Lx, Ly, cx=cx, cy=cy, amp=amp:
                      lens_velocity_rect(nx, ny, Lx, Ly, cx * Lx, cy * Ly, amp))
        else:
            vel_fn = lambda nx, ny, Lx, Ly, s=pseed: layered_velocity(nx, ny, Lx, Ly, s)
        src_fn = lambda Lx, h, sx=sx, sy=sy: [(float(a * Lx), float(b)) for a, b in zip(sx, sy)]
        out.append(_rect_problem(nxn, nyn, (cpy, cpx), vel_fn, ppw, 20, src_fn,
                                 f"p{k}_{'lens' if lens else 'layer'}_{nxn}x{nyn}"))
    return out


SERVE_GRIDS = [(512, 512), (1024, 512), (1024, 1024), (1536, 768), (1600, 800), (2048, 1024), (2400, 800),
               (3000, 1000)]
SERVE_PPW = [6.0, 8.0, 10.0, 12.0]


def config_serve_problem(model, ppw_index, cell_nodes=200):
    """Problem (model, omega) of the serving workload: layered model `model`
    (seed = model) on grid SERVE_GRIDS[model] with cells of about
    `cell_nodes` nodes, SERVE_PPW[ppw_index] points per wavelength, 20 PML
    cells."""
    nxn, nyn = SERVE_GRIDS[model]
    cpx, cpy = max(1, round(nxn / cell_nodes)), max(1, round(nyn / cell_nodes))
    nxn -= nxn % cpx
    nyn -= nyn % cpy
    ppw = SERVE_PPW[ppw_index]
    vel_fn = lambda nx, ny, Lx, Ly, s=model: layered_velocity(nx, ny, Lx, Ly, s)
    return _rect_problem(nxn, nyn, (cpy, cpx), vel_fn, ppw, 20, lambda Lx, h: [],
                         f"m{model}_ppw{int(ppw)}_{nxn}x{nyn}")


def serve_requests(count, rate, seed=0, models=8, freqs=4, max_sources=64):
    """Poisson request stream: exponential inter-arrival times of mean 1/rate
    seconds; each request picks (model, frequency) uniformly and carries 1-64
    sources at random surface positions (x uniform in the interior, 10 nodes
    below the top).  Returns [(t_arrival, model, ppw_index, [(x_frac, depth_nodes)])]."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for _ in range(count):
        t += float(rng.exponential(1.0 / rate))
        m = int(rng.integers(0, models))
        f = int(rng.integers(0, freqs))
        ns = int(rng.integers(1, max_sources + 1))
        xf = rng.uniform(0.05, 0.95, ns)
        out.append((t, m, f, [float(a) for a in xf]))
    return out


CONFIGS = {"lens1600": config_lens_1600, "const400": config_const_400, "layered": config_layered_3x1}
