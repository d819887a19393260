"""Scenario sets for ensembles: seeded tsunami-like uplifts over a common ocean bathymetry.

The ocean is 4000 m deep with a linear shelf rising to 50 m over the last 20% of x (still-water
level 4000, b = 4000 - depth); the boundary is open over the deep ocean and a wall over the shelf
(OceanBoundary.okl with h_zero = 4000).  Each scenario draws its mesh size, degree and an
elliptic Gaussian sea-surface uplift.  Scenarios with the same (size, degree) share one mesh.
"""
from __future__ import annotations

import numpy as np

from .setup_api import Case, RunParams

LENGTH = 800e3          # square domain [0, 800 km]^2
SIZES = (64, 96, 128, 192, 256)
DEGREES = (4, 5, 7)


def uplift_state(case: Case, amp: float, axis_a: float, axis_c: float, angle: float,
                 centre=(0.4, 0.5), length: float = LENGTH) -> np.ndarray:
    """h = 4000 - b + amp exp(-(xr/a)^2 - (yr/c)^2), zero momentum, on the nodes of `case` (an
    ocean mesh); (xr, yr) are the coordinates rotated by `angle` about the centre."""
    x, y, b = case["x"], case["y"], case["b"]
    cx, cy = centre[0] * length, centre[1] * length
    cs, sn = np.cos(angle), np.sin(angle)
    xr = cs * (x - cx) + sn * (y - cy)
    yr = -sn * (x - cx) + cs * (y - cy)
    h = (4000.0 - b) + amp * np.exp(-(xr / axis_a) ** 2 - (yr / axis_c) ** 2)
    q = np.zeros((case.Nelem, 3, case.ngl ** 2))
    q[:, 0] = h.reshape(case.Nelem, -1)
    return q.ravel()


def ocean_run_params(**kw) -> RunParams:
    base = dict(num_flux=3, bc_kind=2, h_zero=4000.0, CFL=0.5, T=1.0e30)
    base.update(kw)
    return RunParams(**base)


def draw(n: int, seed: int = 0, sizes=SIZES, degrees=DEGREES):
    """Seeded scenario parameters: list of dicts (size, N, amp, axis_a, axis_c, angle, centre)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(n):
        out.append(dict(size=int(rng.choice(sizes)), N=int(rng.choice(degrees)),
                        amp=float(rng.uniform(0.5, 5.0)), axis_a=float(rng.uniform(20e3, 100e3)),
                        axis_c=float(rng.uniform(20e3, 100e3)), angle=float(rng.uniform(0.0, np.pi)),
                        centre=(float(rng.uniform(0.25, 0.55)), float(rng.uniform(0.3, 0.7)))))
    return out


def build(params, rp: RunParams | None = None):
    """Members (case, rp, q) for an ensemble; one Case per distinct (size, N)."""
    rp = rp or ocean_run_params()
    meshes = {}
    members = []
    for p in params:
        key = (p["size"], p["N"])
        if key not in meshes:
            meshes[key] = Case.ocean_uplift(N=p["N"], nel=p["size"], amp=0.0, length=LENGTH)
        case = meshes[key]
        q = uplift_state(case, p["amp"], p["axis_a"], p["axis_c"], p["angle"], p["centre"])
        members.append((case, rp, q))
    return members
