"""Synthetic workloads of the local stage and a one-call runner.

* ``config1``: 8x8 unit-square macros, each ear-split and refined to 4x4x2 P1
  triangles (25 local DOFs), piecewise-constant multipliers on 2 segments per
  macroface (m = 8), K = 1, f = 2 pi^2 sin(pi x) sin(pi y), homogeneous
  Dirichlet data (zero weak boundary load), quadrature order 4.
* ``config2``: 64x64 macros of 16x16x2 P2 triangles (1089 DOFs), discontinuous
  P1 multipliers on 4 segments per macroface (m = 32), log K a Gaussian random
  field with exponential covariance (correlation length 0.05, variance 4,
  fixed seed) sampled at the 16 quadrature points of every triangle, f = 1.
* ``config3``: nx^3 cube macros of r^3 lattice cubes x 6 Freudenthal P2
  tetrahedra, piecewise-constant multipliers on 2 x 2 cells of every
  macroface (m = 24), K a binary inclusion field (seeded random spheres of
  radius 0.02 covering 20% of the volume, K = 1e-3 inside, 1 outside) at the
  64 quadrature points of every tetrahedron, f = 1.  The full workload is
  nx = 16, r = 4 (n = 729, bandwidth 189, factored by the wide-band kernel
  and solved by the block solve); r = 2 (n = 125, bandwidth 65) is the small
  size that fits the LDS-window factor and the wavefront solve.
* ``config4``: Voronoi macros of 2048 seeded points in the unit square (edges
  under a twentieth of the seed spacing collapsed), each ear-split and refined
  to about h = 1/1024 within 200-3000 P1 DOFs, discontinuous P1 multipliers per
  macroface, log K a Gaussian field (correlation length 0.02, variance 2),
  f = 1.  Macros of equal local topology form buckets (``PolygonProblem``).

Fields are generated here (seeded) because the reference holds no random
field generator for these cases.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import meshing
from .problems import CubeGridProblem, PolygonProblem, SquareGridProblem


@dataclass
class Case:
    name: str
    problem: SquareGridProblem
    kq: np.ndarray            # (M, ne, nq)
    fq: np.ndarray | None     # (M, ne, nq) or None
    f_const: float = 0.0


def gaussian_field(points, corr_len: float, variance: float, seed: int, grid: int = 1024):
    """Stationary Gaussian field with covariance ``variance * exp(-r / corr_len)``.

    Sampled exactly on a periodic ``2 grid x 2 grid`` lattice over [0, 2)^2 by
    circulant embedding (negative embedding eigenvalues clipped to zero), then
    bilinearly interpolated at ``points`` in [0, 1]^2.
    """
    n2 = 2 * grid
    h = 1.0 / grid
    d = np.minimum(np.arange(n2), n2 - np.arange(n2)) * h
    r = np.sqrt(d[:, None] ** 2 + d[None, :] ** 2)
    lam = np.fft.fft2(variance * np.exp(-r / corr_len)).real
    lam = np.maximum(lam, 0.0)
    rng = np.random.default_rng(seed)
    xi = rng.standard_normal((n2, n2)) + 1j * rng.standard_normal((n2, n2))
    field = np.fft.fft2(np.sqrt(lam / (n2 * n2)) * xi).real[: grid + 1, : grid + 1]
    p = np.asarray(points, dtype=float).reshape(-1, 2)
    x = np.clip(p[:, 0] / h, 0, grid - 1e-9)
    y = np.clip(p[:, 1] / h, 0, grid - 1e-9)
    i, j = np.floor(x).astype(np.int64), np.floor(y).astype(np.int64)
    fx, fy = x - i, y - j
    val = ((1 - fx) * (1 - fy) * field[i, j] + fx * (1 - fy) * field[i + 1, j]
           + (1 - fx) * fy * field[i, j + 1] + fx * fy * field[i + 1, j + 1])
    return val.reshape(np.shape(points)[:-1])


def config1(nx: int = 8, device: int = 0) -> Case:
    prob = SquareGridProblem(nx, 4, 1, meshing.FaceSpace.uniform(0, 2), order=4, device=device)
    q = prob.quadrature_points()
    kq = np.ones(q.shape[:3])
    fq = 2 * np.pi ** 2 * np.sin(np.pi * q[..., 0]) * np.sin(np.pi * q[..., 1])
    return Case("config1", prob, kq, fq)


def config2(nx: int = 64, seed: int = 20261203, device: int = 0, refinement: int = 16) -> Case:
    prob = SquareGridProblem(nx, refinement, 2, meshing.FaceSpace.uniform(1, 4), order=4, device=device)
    q = prob.quadrature_points()
    logk = gaussian_field(q, 0.05, 4.0, seed)
    return Case("config2", prob, np.exp(logk), None, 1.0)


def sphere_field(points, radius: float, fraction: float, seed: int, inside: float, outside: float):
    """Binary inclusion field: ``inside`` within any of N seeded uniform random
    spheres in [0, 1]^3 (N from the Boolean-model volume fraction
    1 - exp(-N 4/3 pi r^3) = fraction), ``outside`` elsewhere."""
    from scipy.spatial import cKDTree

    count = int(round(-np.log(1.0 - fraction) / (4.0 / 3.0 * np.pi * radius ** 3)))
    centres = np.random.default_rng(seed).random((count, 3))
    p = np.asarray(points, dtype=float).reshape(-1, 3)
    dist, _ = cKDTree(centres).query(p, distance_upper_bound=radius)
    return np.where(np.isfinite(dist), inside, outside).reshape(np.shape(points)[:-1])


def config3(nx: int = 16, refinement: int = 4, seed: int = 20261204, device: int = 0) -> Case:
    prob = CubeGridProblem(nx, refinement, 2, split=2, order=4, device=device)
    kq = sphere_field(prob.quadrature_points(), 0.02, 0.2, seed, 1e-3, 1.0)
    return Case("config3", prob, kq, None, 1.0)


def polygon_refinements(macro, h: float, dof_range=(200, 3000),
                        levels=(4, 6, 8, 10, 12, 14, 16, 19, 22, 26, 30, 35, 40, 46)):
    """Per-macro ear refinement for a target element size ``h``: the level
    closest to ``sqrt(area) / h`` among those giving ``dof_range`` P1 DOFs for
    the macro's vertex count (a fan of k - 2 ears refined r times has
    (k-2)(r+1)(r+2)/2 - (k-3)(r+1) nodes); a few levels keep the buckets few."""
    out = np.empty(len(macro.cells), dtype=np.int64)
    for c, cell in enumerate(macro.cells):
        xy = macro.points[cell]
        k = len(cell)
        area = 0.5 * float(np.sum(meshing._cross(xy - xy[0], np.roll(xy, -1, axis=0) - xy[0])))
        nodes = {r: (k - 2) * (r + 1) * (r + 2) // 2 - (k - 3) * (r + 1) for r in levels}
        ok = [r for r in levels if dof_range[0] <= nodes[r] <= dof_range[1]] or list(levels)
        target = np.sqrt(area) / h
        out[c] = min(ok, key=lambda r: abs(r - target))
    return out


def config4(npts: int = 2048, seed: int = 20261207, h: float = 1.0 / 1024, dof_range=(200, 3000),
            device: int = 0) -> Case:
    """Voronoi macros of ``npts`` seeded points in the unit square, each ear-split
    and refined to about ``h`` within ``dof_range`` P1 DOFs, discontinuous P1
    multipliers per macroface (m = 2 per side), log K a Gaussian field
    (exponential covariance, correlation length 0.02, variance 2), f = 1."""
    macro = meshing.voronoi_square(npts, seed)
    prob = PolygonProblem(macro, polygon_refinements(macro, h, dof_range), 1, meshing.FaceSpace.uniform(1, 1),
                          order=4, device=device)
    q = prob.quadrature_points()
    flat = np.concatenate([x.reshape(-1, 2) for x in q])
    logk = gaussian_field(flat, 0.02, 2.0, seed + 1)
    kq, off = [], 0
    for x in q:
        size = x.shape[0] * x.shape[1]
        kq.append(np.exp(logk[off:off + size]).reshape(x.shape[:2]))
        off += size
    return Case("config4", prob, kq, None, 1.0)


def run_polygon_stage(case: Case, precision: str = "fp64"):
    """The stage over every bucket of a polygon case; per-macro host lists."""
    prob = case.problem
    out = {k: [None] * prob.M for k in ("blocks", "loads", "source", "lifts", "retained")}
    iters = np.zeros(prob.M, dtype=np.int64)
    for bk in prob.buckets:
        with bk.batch(precision) as b:
            kq = np.stack([case.kq[c] for c in bk.cells])
            b.assemble(kq, None, case.f_const)
            b.factor()
            blocks, loads = b.condense()
            src, lifts, ret = b.responses()
            vals = dict(blocks=blocks.cpu().numpy(), loads=loads.cpu().numpy(), source=src.cpu().numpy(),
                        lifts=lifts.cpu().numpy(), retained=ret.cpu().numpy())
            for r, c in enumerate(bk.cells):
                for key in out:
                    out[key][c] = vals[key][r]
            iters[bk.cells] = b.iterations()
    out["iterations"] = iters
    return out


def run_stage(case: Case, precision: str = "fp64", macros: slice | None = None):
    """Assemble, factor, solve and condense on the GPU; return host arrays."""
    sl = slice(None) if macros is None else macros
    prob = case.problem
    with prob.batch(precision, macros=sl) as batch:
        batch.assemble(case.kq[sl], None if case.fq is None else case.fq[sl], case.f_const)
        batch.factor()
        blocks, loads = batch.condense()
        src, lifts, ret = batch.responses()
        return dict(blocks=blocks.cpu().numpy(), loads=loads.cpu().numpy(), source=src.cpu().numpy(),
                    lifts=lifts.cpu().numpy(), retained=ret.cpu().numpy(), iterations=batch.iterations())
