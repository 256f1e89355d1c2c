"""Airfoil polygons: the GXXX x-paired point format and seeded NACA-4-style sections.

`reorder_points` restates the reference's polygon preparation (pipeline/geometry.py:10-51):
the trailing-edge top point, the upper surface towards the leading edge, the leading edge,
the lower surface back to the trailing edge and the remaining trailing-edge points, with edge
patches 1 = upper, 2 = lower, 3 = trailing edge.  No point is moved or dropped.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np


def read_points(path: str | Path) -> np.ndarray:
    """Read an `x,y` CSV (header row required) into an (n, 2) float64 array."""
    path = Path(path)
    with path.open(encoding="utf-8-sig") as stream:
        if [x.strip().lower() for x in stream.readline().strip().split(",")] != ["x", "y"]:
            raise ValueError("CSV first row must be x,y")
    points = np.loadtxt(path, delimiter=",", skiprows=1, ndmin=2)
    if points.shape[1] != 2 or not np.isfinite(points).all():
        raise ValueError("expected CSV header x,y followed by two finite numeric columns")
    return points


def reorder_points(points: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """GXXX x-paired points -> (counter-clockwise polygon (n, 2), patch id per edge (n,))."""
    points = np.asarray(points, dtype=np.float64)
    if len(np.unique(points, axis=0)) != len(points):
        raise ValueError("duplicate points are not accepted; coordinates are never silently removed")
    xs = np.unique(points[:, 0])
    if len(xs) < 3 or xs[0] != 0 or xs[-1] != 1 or np.sum(points[:, 0] == 0) != 1:
        raise ValueError("GXXX format requires normalized x/c in [0,1], one leading-edge point at x=0")
    upper, lower = [], []
    for x in xs[1:-1]:
        ys = np.sort(points[points[:, 0] == x, 1])
        if len(ys) != 2 or ys[0] >= ys[1]:
            raise ValueError(f"x={x}: expected distinct lower and upper points, exactly two")
        lower.append((x, ys[0]))
        upper.append((x, ys[1]))
    te = points[points[:, 0] == 1]
    te = te[np.argsort(te[:, 1])]
    if len(te) < 2:
        raise ValueError("GXXX format requires a finite-thickness vertical trailing edge with >=2 points")
    leading = points[points[:, 0] == 0][0]
    vertices = np.array([te[-1], *upper[::-1], leading, *lower, *te[:-1]], dtype=np.float64)
    if len(vertices) != len(points) or set(map(tuple, vertices)) != set(map(tuple, points)):
        raise ValueError("polygon reordering did not preserve the complete coordinate set")
    area = float(np.sum(vertices[:, 0] * np.roll(vertices[:, 1], -1) - vertices[:, 1] * np.roll(vertices[:, 0], -1)) / 2)
    if area <= 0:
        raise ValueError("airfoil polygon has nonpositive area")
    patches = np.array([1] * (len(upper) + 1) + [2] * (len(lower) + 1) + [3] * (len(te) - 1), dtype=np.int32)
    return vertices, patches


def naca4_points(thickness: float, camber: float = 0.0, camber_position: float = 0.4,
                 n_stations: int = 100) -> np.ndarray:
    """x-paired NACA-4-digit-style section in GXXX form.

    Stations x_i = (1 - cos(pi i / n)) / 2, i = 0..n (cosine spacing, 2n + 1 points).  The upper
    and lower points share x: y = y_c(x) +/- y_t(x) (camber offset vertical, not normal, so the
    x-pairing survives).  The open NACA thickness law leaves a finite vertical trailing edge
    of height 2 y_t(1) = 0.021 t at x = 1, as the GXXX format requires.
    """
    if not (0 < thickness < 0.5) or not (0 <= camber < 0.2) or not (0.05 <= camber_position <= 0.95):
        raise ValueError("thickness, camber or camber position out of range")
    i = np.arange(n_stations + 1, dtype=np.float64)
    x = 0.5 * (1.0 - np.cos(np.pi * i / n_stations))
    x[0], x[-1] = 0.0, 1.0
    t = thickness
    yt = 5 * t * (0.2969 * np.sqrt(x) - 0.1260 * x - 0.3516 * x**2 + 0.2843 * x**3 - 0.1015 * x**4)
    m, p = camber, camber_position
    if m == 0:
        yc = np.zeros_like(x)
    else:
        yc = np.where(x < p, m / p**2 * (2 * p * x - x**2), m / (1 - p) ** 2 * ((1 - 2 * p) + 2 * p * x - x**2))
    points = [(0.0, float(yc[0]))]
    for k in range(1, n_stations + 1):
        points.append((float(x[k]), float(yc[k] + yt[k])))
        points.append((float(x[k]), float(yc[k] - yt[k])))
    return np.array(points, dtype=np.float64)


def seeded_airfoil(seed: int) -> dict:
    """Batch/serving airfoil: thickness U[.06,.14], camber U[0,.04], position U[.3,.5],
    Mach U[1.2, 2.0], angle of attack U[-2, 4] degrees, from numpy's default_rng(seed)."""
    rng = np.random.default_rng(seed)
    return dict(seed=int(seed), thickness=float(rng.uniform(0.06, 0.14)), camber=float(rng.uniform(0.0, 0.04)),
                camber_position=float(rng.uniform(0.3, 0.5)), mach=float(rng.uniform(1.2, 2.0)),
                aoa=float(rng.uniform(-2.0, 4.0)))


def polygon_area(vertices: np.ndarray) -> float:
    x, y = vertices[:, 0], vertices[:, 1]
    return float(np.sum(x * np.roll(y, -1) - y * np.roll(x, -1)) / 2)


def polygon_perimeter(vertices: np.ndarray) -> float:
    return float(np.linalg.norm(vertices - np.roll(vertices, 1, axis=0), axis=1).sum())


__all__ = ["read_points", "reorder_points", "naca4_points", "seeded_airfoil", "polygon_area",
           "polygon_perimeter"]
