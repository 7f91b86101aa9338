"""Analytic solutions and error measures used to validate the solver (counterpart of validate.py's known answers).

* ``l2`` -- validate.py:50-51.
* ``poiseuille_exact`` -- validate.py:78: u = F / (4 nu) (R^2 - r^2).
* ``womersley_exact`` -- validate.py:86-89: oscillating body force F0 cos(w t), Bessel-function profile.
* ``womersley_flow_profile`` -- the flow-rate-driven Womersley profile for a pipe whose mean velocity is prescribed as
  a Fourier series (used for the pulsatile pipe with a velocity inlet).
* ``womersley_line_flow`` -- how one harmonic of the flow rate varies along a rigid pipe of weakly compressible
  fluid (the lattice fluid, sound speed 1/sqrt(3)) between a velocity inlet and a pressure outlet: Womersley's
  propagation constant for a rigid tube, J. Physiol. 127 (1955) 553-563.
* ``hagen_poiseuille_gradient`` -- validate.py:204: G = 8 nu Q / (pi R^4).
"""
from __future__ import annotations

import math

import numpy as np
from scipy.special import jv, jve


def l2(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


def poiseuille_exact(r, R, F, nu):
    return F / (4 * nu) * (R ** 2 - r ** 2)


def womersley_exact(r, R, Wo, F0, w, t):
    k = (1j ** 1.5) * Wo
    U = F0 / (1j * w) * (1 - jv(0, k * r / R) / jv(0, k))
    return np.real(U * np.exp(1j * w * t))


def womersley_flow_profile(r, R, nu, harmonics, period, t):
    """Axial velocity u(r, t) of fully developed pipe flow whose cross-section mean velocity is
    U(t) = a0 + sum_k a_k cos(k w t) + b_k sin(k w t), w = 2 pi / period (harmonics = [(a0,), (a1, b1), ...]).
    Each harmonic n has the Womersley shape (1 - J0(i^1.5 a r/R) / J0(i^1.5 a)), normalised to unit mean."""
    r = np.asarray(r, float)
    a0 = harmonics[0][0]
    u = 2 * a0 * (1 - (r / R) ** 2)
    w = 2 * math.pi / period
    for n, (a, b) in enumerate(harmonics[1:], start=1):
        Wo = R * math.sqrt(n * w / nu)
        k = (1j ** 1.5) * Wo
        shape = 1 - jv(0, k * r / R) / jv(0, k)
        mean = 1 - 2 * jv(1, k) / (k * jv(0, k))          # cross-section mean of the shape
        amp = complex(a, -b)                               # a cos + b sin = Re[(a - i b) e^{i w t}]
        u = u + np.real(amp * shape / mean * np.exp(1j * n * w * t))
    return u


def womersley_line_flow(z, L, R, nu, period, n=1, c=1 / math.sqrt(3)):
    """Q_n(z) / Q_n(0) for harmonic n of the flow rate in a rigid pipe of length L filled with a fluid of sound speed
    c, the flow rate imposed at z = 0 and the pressure held at z = L (p_n(L) = 0). Linearised, per harmonic
    (w = 2 pi n / period): continuity  i w A p / (rho c^2) = -dQ/dz  and Womersley's momentum balance for the
    cross-section flow  i w rho Q = -A (1 - F10) dp/dz,  F10 = 2 J1(K) / (K J0(K)),  K = i^1.5 R sqrt(w / nu).
    Together  Q'' + k^2 Q = 0  with  k = w / (c sqrt(1 - F10)),  Q(0) = Q0,  Q'(L) = 0, so
    Q(z) = Q0 cos(k (L - z)) / cos(k L). The pipe resonates where Re(k) L nears pi / 2."""
    w = 2 * math.pi * n / period
    K = (1j ** 1.5) * R * math.sqrt(w / nu)
    F10 = 2 * jve(1, K) / (K * jve(0, K))           # scaled Bessel functions: the same ratio, no overflow
    k = w / (c * np.sqrt(1 - F10))
    return np.cos(k * (L - np.asarray(z, float))) / np.cos(k * L)


def hagen_poiseuille_gradient(nu, Q, R):
    return 8 * nu * Q / (math.pi * R ** 4)


# ---------------------------------------------------------------------------------------------------------------
# harness of the validation runs (validate.py:34-47, 151-215), with the solver supplied by the caller
# ---------------------------------------------------------------------------------------------------------------
RHO_BLOOD = 1060.0          # kg/m^3 (vessel.py:46)
MU_BLOOD = 3.5e-3           # Pa s (vessel.py:47)
NU_BLOOD = MU_BLOOD / RHO_BLOOD
OPERATING = {"healthy": (1.0e-4, 5.0e-6), "narrow70": (1.0e-4, 2.0e-6), "narrow70_fine": (5.0e-5, 1.0e-6)}
SPONGE_TAU = 0.55


def units(key):
    from .lattice import Units
    dx, dt = OPERATING[key]
    return Units(dx, dt, RHO_BLOOD, NU_BLOOD)


def section_profile(sc, values, L, nb=None):
    """Mean of `values` in 1-cell slabs along the axis -> (slab centres, means) (validate.py:151-158)."""
    nb = nb or int(L)
    edges = np.linspace(0, L, nb + 1)
    k = np.clip(np.digitize(sc, edges) - 1, 0, nb - 1)
    cnt = np.bincount(k, minlength=nb)
    m = np.bincount(k, weights=values, minlength=nb) / np.maximum(cnt, 1)
    return 0.5 * (edges[1:] + edges[:-1]), m


def flow_rate(sc, u_axial, s0, s1):
    """Volume flow through the pipe averaged over the slab s0..s1 (validate.py:161-164)."""
    m = (sc >= s0) & (sc < s1)
    return float(u_axial[m].sum() / (s1 - s0))


def run_steady(s, measure, chunk, min_chunks=4, max_chunks=40, tol=2e-4):
    """Run chunks until the first measured quantity changes by <= tol relative (validate.py:167-176)."""
    hist = []
    for it in range(max_chunks):
        s.run(chunk)
        if not s.ok():
            return hist, False
        hist.append(measure())
        if it + 1 >= min_chunks and abs(hist[-1][0] - hist[-2][0]) <= tol * abs(hist[-1][0]):
            break
    return hist, True


def open_pipe_case(make_solver, axis, D=20, length_d=12, U_phys=0.6, key="healthy", chunk=10000, min_chunks=6,
                   max_chunks=40):
    """Open pipe with velocity inlet and pressure outlet at the artery's operating point (validate.py:179-215).
    make_solver(lat, nu_lu) returns a solver with the lbm.Solver interface."""
    from .geometry import open_pipe
    un = units(key)
    U = float(un.u_lu(U_phys))
    R = D / 2
    L = length_d * D
    lat, sc, rc, ax = open_pipe(axis, L, lambda s: np.full_like(s, R, dtype=float))
    s = make_solver(lat, un.nu_lu)
    s.set_sponge(D, SPONGE_TAU)
    s.set_drive([U], ramp=4000)
    s.set_outlet_pressure([0.0])
    a, b = 3 * D, (length_d - 3) * D

    def measure():
        p, u = s.fields()
        ua = u @ ax
        Q = flow_rate(sc, ua, a, b)
        sm, pm = section_profile(sc, p, L)
        m = (sm > a) & (sm < b)
        slope = np.polyfit(sm[m], pm[m], 1)[0]
        return -slope, Q, sm, pm, ua

    hist, ok = run_steady(s, measure, chunk=chunk, min_chunks=min_chunks, max_chunks=max_chunks)
    G, Q, sm, pm, ua = hist[-1]
    G_exact = hagen_poiseuille_gradient(un.nu_lu, Q, R)
    mid = (sc > a) & (sc < b)
    Um = Q / (math.pi * R * R)
    prof_err = l2(ua[mid], 2 * Um * (1 - (rc[mid] / R) ** 2))
    return {"cells": lat["nbr"].shape[1], "steps": s.steps, "ok": ok, "grad_err": G / G_exact - 1,
            "flow_vs_inlet": Q / (U * math.pi * R * R) - 1, "profile_l2": prof_err, "G": G, "G_exact": G_exact,
            "counts": lat["counts"]}
