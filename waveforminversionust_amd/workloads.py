"""Synthetic workloads: grids, transducer rings and sound-speed models.

These build the inputs of the documented benchmark cases (there is no dataset:
every model is generated from a seed) and the small fixtures of the
reference's own tests:

* ``ring_case(...)``           -- square grid with a PML band, a centred ring
  of point transducers (every element transmits, all receive), and a model.
* ``disc_model`` / ``smooth_random_model`` -- the two model families.
* ``runtime_test_case()``      -- the 31x31 fixture of
  Runtime/tests/test_runtime.m:10-28 / Runtime/tests/wust_test_measurements.m:14.
* ``probe_case()``             -- the 7x7 probe of Runtime/matlab/wust_probe.m:17-22.
* ``continuation_cases()``     -- eight frequency stages 0.3-1.0 MHz, each on its own grid at
  6 points per wavelength (201^2 ... 669^2 interiors + 24-cell PML), 256 TX each.
* ``breast_case()``            -- 1088^2 nodes, 512-element ring, 12 frequencies, breast phantom.

Geometry follows Runtime/matlab/wust_geometry.m: pixel-edge origin, node
coordinates at pixel centres, nearest-node snapping with the lowest index on
ties, 1-based MATLAB linear indices ``y + Ny*x + 1``.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class Case:
    """One Helmholtz inversion case (one model, one or more frequencies)."""

    name: str
    ny: int
    nx: int
    dx: float
    x_m: np.ndarray
    y_m: np.ndarray
    tx_index: np.ndarray  # 1-based linear node indices
    rx_index: np.ndarray
    c_true: np.ndarray    # [ny, nx] m/s
    frequencies_hz: list
    cfg: dict = field(default_factory=dict)
    tx_xy: np.ndarray | None = None
    rx_xy: np.ndarray | None = None

    @property
    def n_tx(self):
        return int(self.tx_index.size)

    @property
    def n_rx(self):
        return int(self.rx_index.size)


def grid_axes(n, dx):
    """Centred pixel-edge grid of n nodes: coordinates (i + 0.5) dx - n dx / 2."""
    origin = -n * dx / 2.0
    return origin + (np.arange(n) + 0.5) * dx, origin


def snap(xy, x, y):
    """Nearest-node 1-based linear indices (Runtime/matlab/wust_geometry.m:21-31)."""
    xy = np.asarray(xy, np.float64)
    if xy.ndim != 2 or xy.shape[1] != 2 or not np.all(np.isfinite(xy)):
        raise ValueError("WUST:InvalidInput: positions must be finite [N,2] physical [x,y]")
    if not (np.all((xy[:, 0] >= x[0]) & (xy[:, 0] <= x[-1]) & (xy[:, 1] >= y[0]) & (xy[:, 1] <= y[-1]))):
        raise ValueError("WUST:InvalidInput: array outside grid sampling coordinates")
    ix = np.argmin(np.abs(x[:, None] - xy[None, :, 0]), axis=0)
    iy = np.argmin(np.abs(y[:, None] - xy[None, :, 1]), axis=0)
    idx = iy + y.size * ix + 1
    if np.unique(idx).size != idx.size:
        raise ValueError("WUST:InvalidInput: distinct elements in one array collapse onto a grid sample")
    return idx.astype(np.int64)


def ring(n_elem, radius):
    """Positions [x, y] of n_elem points equally spaced on a centred circle."""
    th = 2 * np.pi * np.arange(n_elem) / n_elem
    return np.stack([radius * np.cos(th), radius * np.sin(th)], axis=1)


def disc_model(x, y, background, discs):
    """Background speed plus hard discs given as (speed, radius, cx, cy)."""
    X, Y = np.meshgrid(x, y)
    c = np.full(X.shape, float(background))
    for speed, r, cx, cy in discs:
        c[(X - cx) ** 2 + (Y - cy) ** 2 < r * r] = speed
    return c


def gaussian_filter2(a, sigma_cells):
    """Separable Gaussian filter (reflecting edges, truncated at 4 sigma)."""
    rad = int(np.ceil(4 * sigma_cells))
    t = np.arange(-rad, rad + 1)
    k = np.exp(-0.5 * (t / sigma_cells) ** 2)
    k /= k.sum()
    out = a
    for axis in (0, 1):
        padw = [(0, 0), (0, 0)]
        padw[axis] = (rad, rad)
        p = np.pad(out, padw, mode="reflect")
        acc = np.zeros_like(out)
        for i, w in enumerate(k):
            sl = [slice(None), slice(None)]
            sl[axis] = slice(i, i + out.shape[axis])
            acc += w * p[tuple(sl)]
        out = acc
    return out


def smooth_random_model(ny, nx, dx, seed, amplitude=40.0, sigma_m=3e-3, background=1500.0):
    """1500 m/s plus Gaussian-filtered seeded noise scaled to +-amplitude."""
    rng = np.random.default_rng(seed)
    g = gaussian_filter2(rng.standard_normal((ny, nx)), sigma_m / dx)
    return background + amplitude * g / np.max(np.abs(g))


def ring_case(name, interior, pml_cells, dx, n_elem, radius, freqs, model, seed=0):
    n = interior + 2 * pml_cells
    x, _ = grid_axes(n, dx)
    y = x.copy()
    pos = ring(n_elem, radius)
    tx = snap(pos, x, y)
    if model == "discs":
        c = disc_model(x, y, 1500.0, [(1550.0, 5e-3, 8e-3, 0.0), (1470.0, 4e-3, -6e-3, 5e-3)])
    else:
        c = smooth_random_model(n, n, dx, seed)
    cfg = dict(backend="gpu", wavenumber="continuum", pml_strength=10.0, pml_m=pml_cells * dx,
               stencil_bounds=[1400.0, 1700.0], source_batch_size=n_elem)
    return Case(name=name, ny=n, nx=n, dx=dx, x_m=x, y_m=y, tx_index=tx, rx_index=tx.copy(),
                c_true=c, frequencies_hz=list(freqs), cfg=cfg, tx_xy=pos, rx_xy=pos.copy())


def cpu_reference_case():
    """128x128 interior + 16-cell PML, dx 0.5 mm, 0.5 MHz, 64-element ring at 28 mm, two discs."""
    return ring_case("ring160_64tx_0.5MHz", 128, 16, 0.5e-3, 64, 28e-3, [0.5e6], "discs")


def metric_case(seed=20241):
    """512x512 interior + 24-cell PML, dx 0.25 mm, 0.8 MHz, 256-element ring at 56 mm, smooth model."""
    return ring_case("ring560_256tx_0.8MHz", 512, 24, 0.25e-3, 256, 56e-3, [0.8e6], "smooth", seed)


def runtime_test_case():
    """31x31 grid at 1 mm, 3 TX / 4 RX, f in {100, 120} kHz (Runtime/tests/test_runtime.m:10-28)."""
    x = np.arange(-15, 16) * 1e-3
    gx = -0.0155 + (np.arange(31) + 0.5) * 1e-3
    tx_xy = np.stack([x[[7, 22, 24]], x[[9, 21, 15]]], axis=1)
    rx_xy = np.stack([x[[12, 23, 14, 7]], x[[7, 14, 22, 17]]], axis=1)
    tx = snap(tx_xy, gx, gx)
    rx = snap(rx_xy, gx, gx)
    X, Y = np.meshgrid(gx, gx)
    c = 1500 + 30 * np.exp(-(X**2 + Y**2) / 3e-5)
    cfg = dict(backend="gpu", wavenumber="continuum", pml_strength=10.0, pml_m=0.004,
               stencil_bounds=[1400.0, 1700.0], source_batch_size=2)
    return Case(name="runtime31", ny=31, nx=31, dx=1e-3, x_m=gx, y_m=gx.copy(), tx_index=tx,
                rx_index=rx, c_true=c, frequencies_hz=[1e5, 1.2e5], cfg=cfg, tx_xy=tx_xy, rx_xy=rx_xy)


def continuation_cases(n_stages=8, f_lo=0.3e6, f_hi=1.0e6, domain_m=0.156, pml_cells=24, n_elem=256,
                       radius=70e-3, points_per_wavelength=6, c_min=1400.0):
    """Frequency-continuation stages, each on its own grid: dx = c_min / (ppw f), a domain_m-wide
    interior plus a pml_cells PML, the same 256-element ring and two-disc model sampled per grid.
    With the defaults the interiors run from 201^2 (0.3 MHz, dx 0.78 mm) to 669^2 (1.0 MHz, dx 0.23 mm)."""
    out = []
    for f in np.linspace(f_lo, f_hi, n_stages):
        dx = c_min / (points_per_wavelength * f)
        interior = int(round(domain_m / dx))
        out.append(ring_case(f"stage_{f / 1e6:.1f}MHz_{interior + 2 * pml_cells}", interior, pml_cells, dx, n_elem,
                             radius, [float(f)], "discs"))
    return out


def breast_phantom(x, y, seed=7, n_inclusions=6):
    """1480 m/s fat, a 1540 m/s fibroglandular ellipse (40 x 30 mm), and seeded round inclusions
    of 1450-1580 m/s with radii 2-6 mm placed inside the ellipse's neighbourhood (r < 35 mm)."""
    X, Y = np.meshgrid(x, y)
    c = np.full(X.shape, 1480.0)
    c[(X / 20e-3) ** 2 + (Y / 15e-3) ** 2 <= 1] = 1540.0
    rng = np.random.default_rng(seed)
    for _ in range(n_inclusions):
        r = rng.uniform(2e-3, 6e-3)
        rho, th = 35e-3 * np.sqrt(rng.uniform()), rng.uniform(0, 2 * np.pi)
        c[(X - rho * np.cos(th)) ** 2 + (Y - rho * np.sin(th)) ** 2 <= r * r] = rng.uniform(1450.0, 1580.0)
    return c


def breast_case(seed=7):
    """1024^2 interior + 32-cell PML (1088^2 nodes), dx 0.125 mm, 512-element ring at 60 mm,
    12 frequencies 0.3-1.2 MHz, breast phantom."""
    n, pml, dx = 1024, 32, 0.125e-3
    x, _ = grid_axes(n + 2 * pml, dx)
    pos = ring(512, 60e-3)
    tx = snap(pos, x, x)
    cfg = dict(backend="gpu", wavenumber="continuum", pml_strength=10.0, pml_m=pml * dx,
               stencil_bounds=[1400.0, 1700.0], source_batch_size=512)
    return Case(name="breast1088_512tx_12f", ny=n + 2 * pml, nx=n + 2 * pml, dx=dx, x_m=x, y_m=x.copy(),
                tx_index=tx, rx_index=tx.copy(), c_true=breast_phantom(x, x, seed),
                frequencies_hz=[float(f) for f in np.linspace(0.3e6, 1.2e6, 12)], cfg=cfg, tx_xy=pos,
                rx_xy=pos.copy())


def probe_case():
    """7x7 homogeneous probe with a centre source (Runtime/matlab/wust_probe.m:17-22)."""
    x = np.arange(-3, 4) * 1e-3
    c = 1500.0 * np.ones((7, 7))
    cfg = dict(backend="gpu", wavenumber="continuum", pml_strength=10.0, pml_m=0.001,
               stencil_bounds=[1400.0, 1700.0], source_batch_size=1)
    tx = np.array([3 + 7 * 3 + 1], np.int64)
    return Case(name="probe7", ny=7, nx=7, dx=1e-3, x_m=x, y_m=x.copy(), tx_index=tx, rx_index=tx.copy(),
                c_true=c, frequencies_hz=[1e5], cfg=cfg)
