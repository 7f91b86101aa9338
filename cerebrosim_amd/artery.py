"""Counterpart of simulate.py's solver-side machinery: resistance outlets and per-step flow measurement captured in
the solver's graphs, probe reductions, the lumped circuit that sets the inlet pressure waveform, and the unit
mapping of a run.

* ``ArterySolver`` -- simulate.py:93-114: after each step, the flow through each opening's slab (k_flux) and the
  opening pressures (k_bc: inlet pressure = drive, outlets relaxed towards R Q) -- one deterministic HIP kernel,
  ``cs_flux_bc``, appended to ``_launch`` so it is captured into the graphs.
* ``ArterySolver.gather`` -- simulate.py:85-90, 305-318: probe mean pressures and the maximum speed (``cs_gather``).
* ``INFLOW_HARMONICS`` / ``inflow_waveform`` -- the published inflow waveform (geometry/physics.json "inlet",
  generated by vessel.py:253-265) as data; ``circuit`` -- simulate.py:138-148; ``step_size`` -- simulate.py:155-159;
  ``flux_slabs`` -- simulate.py:180-192.
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
OUTLET_RELAX_S = 0.5e-3       # simulate.py:55
END_LABELS = ("inlet (ICA)", "MCA outlet", "ACA outlet")   # the circuit's openings, in its order (simulate.py:56)


# Inflow over one beat, Q(t) / Q_mean, as the trigonometric series a0 + sum_k a_k cos(2 pi k t / T) + b_k sin(...),
# k = 1..10, coefficients [a0, a1, b1, ..., a10, b10]. Data: the least-squares fit of the reference's published
# waveform (geometry/physics.json "inlet", 200 samples; tests/golden/fit_inflow_harmonics.py), with a0 = 1 (the
# waveform is normalised to its mean). It reproduces the published samples to their rounding (5e-6).
INFLOW_HARMONICS = np.array((
    1.0, 0.069498890, 0.241689468, -0.053443492, 0.104796822, -0.052107328, 0.065175090, -0.071062744,
    0.018654888, -0.050795072, -0.032617016, 0.008394505, -0.043637227, 0.024404578, -0.001922961, 0.009999664,
    0.007282970, 0.004774265, 0.010065523, -0.001034301, 0.006781041))


def series(coef, phase):
    """Value of the trigonometric series coef = [a0, a1, b1, a2, b2, ...] at phase = t / period (array)."""
    coef = np.asarray(coef, float)
    k = np.arange(1, (len(coef) - 1) // 2 + 1)
    z = np.exp(2j * math.pi * np.multiply.outer(np.asarray(phase, float), k))      # (..., harmonics)
    return coef[0] + z.real @ coef[1::2][:len(k)] + z.imag @ coef[2::2][:len(k)]


def inflow_waveform(n=200):
    """(t in s, Q(t) / Q_mean) at n equally spaced times over one beat, from INFLOW_HARMONICS."""
    ph = np.arange(n) / n
    return ph * HEART_PERIOD, series(INFLOW_HARMONICS, ph)


def circuit(R_seg):
    """The lumped circulation around the segment (SI units). Each outlet k is a resistance to the common venous
    pressure P_OUT that passes its share of the mean inflow: R_k = P_OUT / Q_k (MCA: the target flow, ACA: the
    rest), so together they are R_eq = P_OUT / Q_in. The segment adds R_seg in series, and the inlet pressure that
    drives the sourced inflow waveform through both is P_in(t) = (R_eq + R_seg) Q_in(t), returned as the
    coefficients of its trigonometric series (absolute Pa). Counterpart of simulate.py:138-148."""
    q_in = Q_INLET_MEAN / 6e7                         # mL/min -> m^3/s
    q_mca = Q_MCA_TARGET / 6e7
    p_ven = P_OUT_MMHG * MMHG
    r_eq = p_ven / q_in
    return {"R_out": [0.0, p_ven / q_mca, p_ven / (q_in - q_mca)], "labels": list(END_LABELS), "R_seg": R_seg,
            "P_in_coef": ((r_eq + R_seg) * q_in * INFLOW_HARMONICS).tolist(), "R_eq": r_eq}


def outlet_resistances(cal, ends):
    """The circuit's resistance of each open end, matched by the end's label (simulate.py:272-273). A calibration
    written without labels lists them in END_LABELS order. Raises KeyError for an end the circuit does not have."""
    labels = list(cal.get("labels", END_LABELS))
    out = []
    for e in ends:
        if e["label"] not in labels:
            raise KeyError(f"open end {e['label']!r} has no resistance in the calibration (it has {labels})")
        out.append(cal["R_out"][labels.index(e["label"])])
    return out


def step_size(dx, dt, dt_scale=1.0):
    """Time step snapped so that one sample period is a whole number of steps (simulate.py:155-159)."""
    dt = dt * dt_scale
    return SAMPLE_S / round(SAMPLE_S / dt)


def units(dx, dt):
    return Units(dx, dt, RHO, MU / RHO)


def flux_slabs(lat, ends, dx_mm, slab=FLUX_SLAB_MM):
    """Flow-measurement slabs (simulate.py:180-192). For opening k, the cells whose depth behind the opening plane
    (along its inward normal) lies in [slab[0], slab[1]) mm and whose distance from the opening's axis is below
    1.3 radii. Summing u . w over them gives the flow through the slab per unit length: w = n / thickness, with the
    sign that counts inflow positive at the first opening and outflow positive at the others.
    Returns {"idx", "w", "slot"} (cell ids, weights (M, 3), opening index), openings in order."""
    pos = lat["origin"] + lat["coords"] * dx_mm
    thick = (slab[1] - slab[0]) / dx_mm
    parts = []
    for k, e in enumerate(ends):
        n = np.asarray(e["normal"], float)
        off = pos - np.asarray(e["centre"], float)
        inward = -(off @ n)                                    # depth behind the opening plane (mm)
        radial = off + np.outer(inward, n)                     # component across the axis
        sel = np.nonzero((inward >= slab[0]) & (inward < slab[1])
                         & ((radial * radial).sum(1) < (1.3 * e["radius"]) ** 2))[0]
        wk = (n if k else -n) / thick
        parts.append((sel, np.tile(wk, (sel.size, 1)), np.full(sel.size, k)))
    # C order throughout: the arrays go to C and device code as flat buffers
    return {"idx": np.ascontiguousarray(np.concatenate([p[0] for p in parts]), np.int32),
            "w": np.ascontiguousarray(np.concatenate([p[1] for p in parts]), np.float32),
            "slot": np.ascontiguousarray(np.concatenate([p[2] for p in parts]), np.int32)}


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
            self.f_work = torch.zeros(self.n_open * self.f_chunks, dtype=torch.float64, device=dev)
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

    def _stepped_state(self):
        return [self.Qlast] if hasattr(self, "Qlast") else []

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
