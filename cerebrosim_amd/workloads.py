"""Synthetic case sets for the precision, batching and serving forms.

* ``bifurcation_case``: the precision-form workload -- the stenosed bifurcation of geometry.bifurcation with a
  seeded 8-harmonic pulsatile inlet (mean 0.03, peak 0.08), tau 0.52; ``wall_shear_weights`` gives its WSS fit.

* ``stenosed_tube``: straight tube along z with a Gaussian narrowing, a velocity inlet (the reference's parabolic
  profile, lbm.py:322-328) at z = 0 and a pressure outlet at z = L.
* ``tube_specs``: the batching case set -- n tubes drawn from ``default_rng(seed)``: radius U[8, 48], length
  U[128, 960], area reduction U[0, 0.95] (Gaussian profile, sigma = radius), tau U[0.51, 0.9]. The mean inlet
  speed is chosen per tube (``inlet_speed``) so that the throat speed and the Reynolds number stay in the range
  where the regularised collision is stable. About 150 M cells for n = 96.
* ``request_specs``: the serving request stream -- Poisson arrivals from ``default_rng(seed)``; stenosis
  severity U[0, 1] (diameter reduction, cosine narrowing of validate.py:226-239), mean arterial pressure
  U[70, 145] mmHg, grid spacing from {0.2, 0.1, 0.05} mm; each request runs a number of cardiac cycles.
  Geometry: a 3.2 mm tube (0.62 M cells at 0.1 mm, the size of the reference's story artery at that spacing) with
  the narrowing 4 diameters from the inlet. The mapping of pressure to the lattice follows simulate.py's rule
  that higher pressure runs with a smaller time step (dt_scale 1.0 / 0.75 / 0.6 at 95 / 118 / 146 mmHg,
  interpolated linearly, run_bp.sh:20-24); the cycle is time-compressed to ``beat_steps`` x (0.1 mm / dx) /
  dt_scale steps.
"""
from __future__ import annotations

import math

import numpy as np

from .geometry import open_pipe, stenosis_pipe

DT_SCALE_MMHG = ((95.0, 1.0), (118.0, 0.75), (146.0, 0.6))


def stenosed_tube(R, L, area_frac, sigma=None):
    """Lattice of a tube of radius R, length L (cells) with a Gaussian narrowing of area fraction `area_frac`
    removed at L / 2 (radius R (1 - (1 - sqrt(1 - a)) exp(-(s - L/2)^2 / (2 sigma^2))), sigma = R)."""
    sigma = float(R if sigma is None else sigma)
    depth = 1.0 - math.sqrt(max(0.0, 1.0 - area_frac))

    def radius(s):
        s = np.asarray(s, float)
        return R * (1.0 - depth * np.exp(-0.5 * ((s - 0.5 * L) / sigma) ** 2))

    lat, _, _, _ = open_pipe((0, 0, 1), L, radius)
    return lat


def inlet_speed(R, area_frac, tau, u_max=0.03, throat=0.08, re_max=400.0):
    """Mean inlet speed (lattice units) of a tube: at most u_max, a mean throat speed of at most `throat`, and a
    Reynolds number 2 R u / nu of at most re_max."""
    nu = (tau - 0.5) / 3.0
    return float(min(u_max, throat * (1.0 - area_frac), re_max * nu / (2.0 * R)))


def tube_specs(n=96, seed=1):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        R = float(rng.uniform(8, 48))
        L = float(rng.uniform(128, 960))
        a = float(rng.uniform(0.0, 0.95))
        tau = float(rng.uniform(0.51, 0.9))
        out.append({"name": f"tube{k:02d}", "R": R, "L": L, "area_frac": a, "tau": tau,
                    "u_mean": inlet_speed(R, a, tau)})
    return out


def tube_case(spec):
    """(lattice, nu, setup) of a tube spec; setup(solver) sets the drive, the outlet and a sponge."""
    lat = stenosed_tube(spec["R"], spec["L"], spec["area_frac"])
    nu = (spec["tau"] - 0.5) / 3.0

    def setup(s):
        s.set_drive([spec["u_mean"]], period=1, ramp=200)
        s.set_outlet_pressure([0.0])
        s.set_sponge(min(16.0, spec["R"]), 0.7, kinds=("outlet",))

    return lat, nu, setup


def dt_scale(map_mmhg):
    x, y = zip(*DT_SCALE_MMHG)
    return float(np.interp(map_mmhg, x, y))


def request_specs(n, rate_per_s, seed=2, beats=4, beat_steps=2000, grids=(0.2e-3, 0.1e-3, 0.05e-3)):
    """n requests with Poisson arrival times (seconds from the start)."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate_per_s, n))
    out = []
    for k in range(n):
        sev = float(rng.uniform(0.0, 1.0))
        mapp = float(rng.uniform(70.0, 145.0))
        dx = float(grids[int(rng.integers(len(grids)))])
        ds = dt_scale(mapp)
        period = int(round(beat_steps * (0.1e-3 / dx) / ds))
        out.append({"name": f"req{k:03d}", "arrival": float(t[k]), "severity": sev, "map_mmhg": mapp, "dx": dx,
                    "dt_scale": ds, "period": period, "steps": beats * period})
    return out


def request_case(spec, tau=0.55):
    """(lattice, nu, setup) of a serving request: 3.2 mm tube with a cosine narrowing of diameter reduction
    0.8 x severity; pulsatile inflow (the reference's inflow waveform, 10 harmonics) whose mean lattice speed scales
    with pressure x dt_scale and is capped for the throat."""
    from .artery import INFLOW_HARMONICS
    sev = 0.8 * spec["severity"]
    lat, _, _, D = stenosis_pipe(spec["dx"], sev=sev, D0_mm=3.2, Ls_mm=6.4, up_d=4, down_d=18)
    nu = (tau - 0.5) / 3.0
    area = (1.0 - sev) ** 2
    u_mean = min(0.02 * spec["map_mmhg"] / 95.0 * spec["dt_scale"], 0.05 * area)
    coef = (INFLOW_HARMONICS * u_mean).astype(np.float32)

    def setup(s):
        s.set_drive(coef, period=spec["period"], ramp=spec["period"] // 4)
        s.set_outlet_pressure([0.0])
        s.set_sponge(D, 0.7, kinds=("outlet",))

    return lat, nu, setup


def bifurcation_waveform(seed=0, harmonics=8, mean=0.03, peak=0.08):
    """Fourier coefficients [a0, a1, b1, ...] of the precision-form inlet: harmonics drawn from default_rng(seed)
    with a 1/k fall-off, scaled so that the waveform's mean is `mean` and its maximum `peak`."""
    rng = np.random.default_rng(seed)
    ab = rng.standard_normal((harmonics, 2)) / np.arange(1, harmonics + 1)[:, None]
    t = np.linspace(0.0, 1.0, 2048, endpoint=False)
    g = sum(ab[k, 0] * np.cos(2 * np.pi * (k + 1) * t) + ab[k, 1] * np.sin(2 * np.pi * (k + 1) * t)
            for k in range(harmonics))
    amp = (peak - mean) / g.max()
    coef = [mean] + [float(v) for k in range(harmonics) for v in amp * ab[k]]
    return np.asarray(coef, np.float32)


def bifurcation_case(scale=1.0, tau=0.52, period=None, ramp=None, mean=0.03, peak=0.08):
    """(lattice, nu, setup, info) of the precision-form workload: the stenosed bifurcation (geometry.bifurcation)
    with the seeded 8-harmonic pulsatile inlet, pressure outlets at 0 and sponges at the outlets."""
    from .geometry import bifurcation
    lat, info = bifurcation(scale)
    nu = (tau - 0.5) / 3.0
    period = int(period or round(4000 * scale))
    ramp = int(ramp if ramp is not None else period // 4)
    coef = bifurcation_waveform(0, 8, mean, peak)

    def setup(s):
        s.set_drive(coef, period=period, ramp=ramp)
        s.set_outlet_pressure([0.0, 0.0])
        s.set_sponge(max(4.0, 20.0 * scale), 0.6, kinds=("outlet",))

    return lat, nu, setup, info


def wall_shear_weights(lat, info, dx_mm=0.1, spacing=2.0):
    """Wall points of a bifurcation and their WSS fit weights (wss.weights) on a grid of spacing dx_mm:
    (idx (V, K), c (V, K), normals (V, 3), points (V, 3))."""
    from .geometry import wall_points
    from .wss import weights
    P, Nn = wall_points(info, spacing=spacing)
    cells = lat["coords"].astype(float)
    depth = -info["sdf"](cells) * dx_mm
    idx, c, _, _ = weights(P * dx_mm, Nn, cells * dx_mm, depth)
    return idx, c, Nn.astype(np.float32), P
