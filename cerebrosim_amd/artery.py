"""Counterpart of simulate.py's solver-side machinery: resistance outlets and per-step flow measurement captured in
the solver's graphs, probe reductions, the lumped circuit that sets the inlet pressure waveform, and the unit
mapping of a run.

* ``ArterySolver`` -- simulate.py:93-114: after each step, the flow through each opening's slab (k_flux) and the
  opening pressures (k_bc: inlet pressure = drive, outlets relaxed towards R Q) -- one deterministic HIP kernel,
  ``cs_flux_bc``, appended to ``_launch`` so it is captured into the graphs.
* ``ArterySolver.gather`` -- simulate.py:85-90, 305-318: probe mean pressures and the maximum speed (``cs_gather``).
* ``inflow_waveform`` -- vessel.py:253-265; ``fourier`` / ``from_fourier`` -- simulate.py:122-135;
  ``circuit`` -- simulate.py:138-148; ``step_size`` -- simulate.py:155-159; ``flux_slabs`` -- simulate.py:180-192.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _native
from .lbm import Solver, Units

MMHG = 133.322
RHO = 1060.0                  # kg/m^3 (vessel.py:46)
MU = 3.5e-3                   # Pa s (vessel.py:47)
HEART_PERIOD = 1.0            # s (vessel.py:48)
Q_INLET_MEAN = 240.0          # mL/min (vessel.py:49)
Q_MCA_TARGET = 135.0          # mL/min (vessel.py:50)
PULSATILITY = 0.83            # (vessel.py:51)
P_OUT_MMHG = 80.0             # simulate.py:48
SPONGE_MM, SPONGE_TAU = 2.0, 0.55       # simulate.py:50
FLUX_SLAB_MM = (2.5, 3.5)     # simulate.py:51
SAMPLE_S = 1.0e-3             # simulate.py:52
HARMONICS = 10                # simulate.py:54
OUTLET_RELAX_S = 0.5e-3       # simulate.py:55


def inflow_waveform(n=200):
    """Inflow over one beat Q(t) / Q_mean (mean 1, (max - min) / mean = PULSATILITY): systolic peak, late-systolic
    shoulder and diastolic decay, low-passed to 10 harmonics (vessel.py:253-265)."""
    t = np.arange(n) / n
    g = lambda c, w: np.exp(-0.5 * ((t - c) / w) ** 2)
    shape = 0.15 + 1.0 * g(0.13, 0.045) + 0.45 * g(0.30, 0.06) + 0.12 * g(0.45, 0.04) + 0.25 * np.exp(-(t / 0.9))
    F = np.fft.rfft(shape)
    F[11:] = 0
    shape = np.fft.irfft(F, n)
    shape = shape / shape.mean()
    shape = 1 + (shape - 1) * PULSATILITY / (shape.max() - shape.min())
    return t * HEART_PERIOD, shape


def fourier(y, nh=HARMONICS):
    """Samples over one period -> [a0, a1, b1, a2, b2, ...] for a0 + sum a_k cos(2 pi k t) + b_k sin(2 pi k t)."""
    F = np.fft.rfft(y) / len(y)
    c = [F[0].real]
    for k in range(1, nh + 1):
        c += [2 * F[k].real, -2 * F[k].imag]
    return np.array(c)


def from_fourier(c, t):
    y = np.full_like(np.asarray(t, float), c[0], dtype=float)
    for k in range(1, (len(c) - 1) // 2 + 1):
        y += c[2 * k - 1] * np.cos(2 * math.pi * k * t) + c[2 * k] * np.sin(2 * math.pi * k * t)
    return y


def circuit(R_seg):
    """Lumped circuit (SI): outlet resistances and the inlet pressure waveform (Fourier coefficients, Pa absolute)
    that draws the mean inflow through segment + outlets (simulate.py:138-148)."""
    q_in = Q_INLET_MEAN * 1e-6 / 60
    q_mca = Q_MCA_TARGET * 1e-6 / 60
    p_out = P_OUT_MMHG * MMHG
    R_mca, R_aca = p_out / q_mca, p_out / (q_in - q_mca)
    R_eq = p_out / q_in
    _, shape = inflow_waveform(400)
    coef = fourier((R_eq + R_seg) * q_in * shape)
    return {"R_out": [0.0, R_mca, R_aca], "R_seg": R_seg, "P_in_coef": coef.tolist(), "R_eq": R_eq}


def step_size(dx, dt, dt_scale=1.0):
    """Time step snapped so that one sample period is a whole number of steps (simulate.py:155-159)."""
    dt = dt * dt_scale
    return SAMPLE_S / round(SAMPLE_S / dt)


def units(dx, dt):
    return Units(dx, dt, RHO, MU / RHO)


def flux_slabs(lat, ends, dx_mm, slab=FLUX_SLAB_MM):
    """Cells of each opening's measurement slab and their weights +-normal / thickness; inflow positive at the
    first opening, outflow positive at the others (simulate.py:180-192)."""
    P = lat["origin"] + lat["coords"] * dx_mm
    idx, w, slot = [], [], []
    thick = (slab[1] - slab[0]) / dx_mm
    for k, e in enumerate(ends):
        normal = np.asarray(e["normal"], float)
        rel = P - np.asarray(e["centre"], float)
        depth = -(rel @ normal)
        side = np.linalg.norm(rel + depth[:, None] * normal, axis=1) < 1.3 * e["radius"]
        m = np.flatnonzero(side & (depth >= slab[0]) & (depth < slab[1]))
        sign = -1.0 if k == 0 else 1.0
        idx.append(m)
        slot.append(np.full(len(m), k))
        w.append(np.tile(sign * normal / thick, (len(m), 1)))
    return {"idx": np.concatenate(idx).astype(np.int32), "w": np.vstack(w).astype(np.float32),
            "slot": np.concatenate(slot).astype(np.int32)}


class ArterySolver(Solver):
    """Solver + resistance outlets and per-step flow measurement inside each step (simulate.py:93-114).

    flux: {"idx" (M,) reference cell ids, "w" (M, 3), "slot" (M,) opening index}; R_lu: resistance per opening in
    lattice units (entry 0, the inlet, unused); pref_lu: reference pressure; alpha: relaxation factor dt / tau_R.
    """

    def __init__(self, lat, nu_lu, flux, R_lu, pref_lu, alpha, **kw):
        super().__init__(lat, nu_lu, **kw)
        import torch
        dev = self.dev
        with torch.cuda.stream(self._tstream):
            self.f_idx = torch.from_numpy(self.backend.slots(flux["idx"])).to(dev)
            self.f_w = torch.from_numpy(np.ascontiguousarray(flux["w"], np.float32)).to(dev)
            self.f_slot = torch.from_numpy(np.ascontiguousarray(flux["slot"], np.int32)).to(dev)
            self.n_open = len(R_lu)
            self.Qlast = torch.zeros(self.n_open, dtype=torch.float32, device=dev)
            self.R = torch.from_numpy(np.asarray(R_lu, np.float32)).to(dev)
            self.f_chunks = max(1, _native.flux_chunks(int(self.f_idx.numel())))
            self.f_work = torch.zeros(self.n_open * self.f_chunks, dtype=torch.float32, device=dev)
        self.synchronize()
        self.pref = float(pref_lu)
        self.alpha = float(alpha)

    def _launch(self, a):
        super()._launch(a)
        elem, comp = self.backend.velocity_addressing()
        A = _native.FluxArgs()
        A.idx, A.w, A.slot, A.m = self.f_idx.data_ptr(), self.f_w.data_ptr(), self.f_slot.data_ptr(), self.f_idx.numel()
        A.vel, A.vel_elem, A.vel_comp = self.backend.velocity(1 - a).data_ptr(), elem, comp
        A.R, A.drive, A.p_out, A.q_last = self.R.data_ptr(), self._drive_ptr(), self.p_out.data_ptr(), \
            self.Qlast.data_ptr()
        A.n_open, A.pref, A.alpha = self.n_open, self.pref, self.alpha
        A.work, A.work_chunks = self.f_work.data_ptr(), self.f_chunks
        _native.call("cs_flux_bc", ctypes.byref(A), self.stream_handle)

    def gather(self, idx, slot, n_probe):
        """Probe reduction on the newest fields: (sum of rho / 3 per probe, max |u| over the probe cells)."""
        import torch
        elem, comp = self.backend.velocity_addressing()
        with torch.cuda.stream(self._tstream):
            gi = torch.from_numpy(self.backend.slots(idx)).to(self.dev)
            gs = torch.from_numpy(np.ascontiguousarray(slot, np.int32)).to(self.dev)
            out = torch.zeros(max(1, n_probe), dtype=torch.float32, device=self.dev)
            umax = torch.zeros(1, dtype=torch.float32, device=self.dev)
        A = _native.GatherArgs()
        A.idx, A.slot, A.m = gi.data_ptr(), gs.data_ptr(), gi.numel()
        A.rho, A.vel, A.vel_elem, A.vel_comp = self.backend.density(self.cur).data_ptr(), \
            self.backend.velocity(self.cur).data_ptr(), elem, comp
        A.rho_elem = self.backend.density_stride()
        A.out, A.umax, A.n_probe = out.data_ptr(), umax.data_ptr(), n_probe
        _native.call("cs_gather", ctypes.byref(A), self.stream_handle)
        self.synchronize()
        return out[:n_probe].cpu().numpy(), float(umax.item())

    def flows(self):
        """Flow through each opening's slab in the last step (lattice units)."""
        self.synchronize()
        return self.Qlast.cpu().numpy()

    def opening_pressures(self):
        self.synchronize()
        return self.p_out.cpu().numpy()
