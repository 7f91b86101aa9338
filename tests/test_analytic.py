"""The analytic solutions the pulsatile-pipe checks compare against, pinned by what the mathematics fixes (no
retyped formula): the prescribed cross-section mean by independent quadrature, no-slip, the momentum equation by
finite differences, and the limits that reduce to textbook cases."""
import math

import numpy as np
from scipy.integrate import quad

from cerebrosim_amd.validation import womersley_flow_profile, womersley_line_flow

R, NU, T = 40.0, 0.02, 4000.0
HARM = [(0.02,), (0.003, 0.01), (-0.001, 0.0005)]


def mean_velocity(t, harm=HARM, period=T, nu=NU):
    w = 2 * math.pi / period
    return harm[0][0] + sum(a * math.cos(k * w * t) + b * math.sin(k * w * t) for k, (a, b) in enumerate(harm[1:], 1))


def test_flow_profile_mean_and_no_slip():
    for t in (0.0, 517.5, 1999.5, 3333.0):
        disc = quad(lambda r: 2 * math.pi * r * float(womersley_flow_profile(r, R, NU, HARM, T, t)), 0, R,
                    epsabs=1e-13, limit=200)[0] / (math.pi * R * R)
        assert abs(disc - mean_velocity(t)) < 1e-9
        assert abs(float(womersley_flow_profile(R, R, NU, HARM, T, t))) < 1e-12


def test_flow_profile_solves_the_momentum_equation():
    """du/dt - nu (u_rr + u_r / r) is the (uniform) pressure gradient: the same at every radius."""
    r = np.linspace(4.0, 36.0, 9)
    h, dt = 0.05, 0.5
    for t in (100.0, 1700.0):
        u = lambda rr, tt: womersley_flow_profile(rr, R, NU, HARM, T, tt)
        ut = (u(r, t + dt) - u(r, t - dt)) / (2 * dt)
        urr = (u(r + h, t) - 2 * u(r, t) + u(r - h, t)) / h ** 2
        ur = (u(r + h, t) - u(r - h, t)) / (2 * h)
        G = ut - NU * (urr + ur / r)
        assert np.ptp(G) < 1e-4 * np.abs(G).max() + 1e-11, G


def test_flow_profile_quasi_steady_limit():
    """At a Womersley number far below one every harmonic is a Poiseuille parabola of the instantaneous mean."""
    r = np.linspace(0, 1.0, 11)
    t = 1.2e8
    u = womersley_flow_profile(r, 1.0, 1.0, HARM, 4.0e9, t)
    assert np.allclose(u, 2 * mean_velocity(t, period=4.0e9, nu=1.0) * (1 - r ** 2), rtol=1e-6, atol=1e-12)


def test_line_flow_boundary_conditions():
    L = 480.0
    assert abs(womersley_line_flow(0.0, L, R, NU, T) - 1) < 1e-12
    h = 1e-3
    d = (womersley_line_flow(L + h, L, R, NU, T) - womersley_line_flow(L - h, L, R, NU, T)) / (2 * h)
    assert abs(d) < 1e-9                       # pressure node at the outlet: dQ/dz = 0 there


def test_line_flow_limits():
    z = np.linspace(0, 480.0, 7)
    # incompressible limit: the flow rate is the same everywhere
    assert np.allclose(womersley_line_flow(z, 480.0, R, NU, T, c=1e9), 1.0, atol=1e-9)
    # inviscid limit (F10 -> 0): the standing acoustic wave cos(w (L - z) / c) / cos(w L / c)
    w, c = 2 * math.pi / T, 1 / math.sqrt(3)
    exact = np.cos(w * (480.0 - z) / c) / np.cos(w * 480.0 / c)
    assert np.allclose(womersley_line_flow(z, 480.0, R, 1e-14, T), exact, rtol=1e-5)


def test_line_flow_friction_matches_the_profile_mean():
    """1 - F10 is the disc mean of the unit-gradient Womersley shape 1 - J0(K r / R) / J0(K): check the wave number
    the line uses against that mean by quadrature, through k^2 c^2 (1 - F10) = w^2."""
    from scipy.special import jv
    w, c = 2 * math.pi / T, 1 / math.sqrt(3)
    K = (1j ** 1.5) * R * math.sqrt(w / NU)
    re = quad(lambda r: 2 * r * (1 - jv(0, K * r / R) / jv(0, K)).real, 0, R, limit=200)[0] / R ** 2
    im = quad(lambda r: 2 * r * (1 - jv(0, K * r / R) / jv(0, K)).imag, 0, R, limit=200)[0] / R ** 2
    one_minus_F10 = complex(re, im)
    # recover k from the line solution itself: Q(z) = cos(k (L - z)) / cos(k L) gives Q'' = -k^2 Q
    L, z, h = 480.0, 240.0, 0.5
    q = [womersley_line_flow(x, L, R, NU, T) for x in (z - h, z, z + h)]
    k2 = -(q[0] - 2 * q[1] + q[2]) / h ** 2 / q[1]
    assert abs(k2 * c ** 2 * one_minus_F10 / w ** 2 - 1) < 1e-5
