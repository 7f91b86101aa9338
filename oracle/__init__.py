"""ORACLE -- TEST INFRASTRUCTURE ONLY.

CPU restatement of CerebroSim's solver step (lbm.py ``k_clock``/``_ddx``/``_h3``/``k_step``) and of the per-step
reductions of simulate.py (``k_flux``/``k_bc``/``k_gather``), in the reference's (Q, N) indexed layout and loop order,
in FP32 and FP64 (C, oracle/lbm_oracle.c + oracle/step_impl.h). Only ``tests/``, ``__graft_entry__.smoke()`` and
the CPU-baseline leg of ``bench.py`` may import this package, and only as the checker or the timed CPU baseline; the
product path (``cerebrosim_amd``) never imports it.

How each function is pinned (so that a plausible mistake fails a check that is not a retyping of the formula):

* ``orc_clock`` (lbm.py:90-102): closed form -- a numpy FP64 Fourier sum with the cosine ramp
  (tests/test_oracle.py::test_clock_closed_form).
* ``orc_step`` streaming + collision (lbm.py:140-257):
  - the reference's published physics results at its own operating point (geometry/solver.json:20-68):
    Poiseuille L2 / flow error at D = 10, 20 (Bouzidi) and D = 20 (half-way walls), Womersley L2 at Wo = 5.6;
  - invariants: mass (sum of post-collision populations = delta rho) and momentum (= j + F) of every collision,
    the rest state staying exactly zero with walls, inlet and outlet links, zero-th and first moments of ``_h3``;
  - special cases: q = 1/2 Bouzidi equals half-way bounce-back, zero-velocity inlet equals bounce-back, outlet with
    p = 0 and u = 0 equals anti-bounce-back -f;
  - FP32 against FP64 on the same inputs.
  The Galilean term and ``_h3`` vanish in the periodic pipes (u = (0, 0, u_z(r))), so they are pinned separately by
  the tilted open pipe's published pressure-gradient error (geometry/solver.json:92-112, slow test) and by the
  brute-force moment checks of ``_h3``.
* the hybrid blend (sigma < 1, lbm.py:224-229): the closed form of the output's second moment with a linear
  previous-velocity field (exact finite differences, off-diagonal gradients included):
  dr/3 I + u u^T + (1 - om) (sigma P - (1 - sigma) / (3 om) (G + G^T) - diag E)
  (tests/test_oracle.py::test_hybrid_blend_second_moment).
* TRT (lbm.py:241-255):
  - the steady planar channel with half-way bounce-back and a body force, whose discrete solution is the mid-link
    parabola plus the uniform slip (F / nu)(2 magic / 3 - 1/8) (Ginzburg 2008; exact for magic = 3/16), in FP64 to
    1e-12 for four values of magic and two of tau (test_trt_channel_poiseuille_closed_form);
  - the odd third moment sum c_x c_y^2 f, which relaxes with om- and carries the Guo source's odd part and the
    third-order equilibrium term (test_trt_odd_third_moment);
  - the second moment, whose closed form equals the regularised one with the Galilean term (test_trt_second_moment_
    equals_regularised_form); mass, momentum and the rest state as for the regularised collision.
* ``orc_flux_bc`` / ``orc_gather`` (simulate.py:65-90): closed forms on hand-made inputs -- per-opening slab sums,
  the inlet pressure copy, the outlet relaxation p += alpha (R Q - pref - p), the accumulator reset; per-probe sums of
  rho / 3 and the running maximum of |u| (test_flux_bc_closed_form, test_gather_closed_form).
* ``oracle/sdf.py`` (vessel.py:183-207, the vessel signed distance): closed forms on a capped cylinder, a cone and
  the smooth minimum (tests/test_vessel_host.py); the lattice built from the checked HIP kernel reproduces the
  published artery counts (geometry/solver.json:179-194).
* ``oracle/wss.py`` (wss.py:35-57): the reference's own known answers, 2 mu u_max / R and 4 mu u_max / R
  (tests/test_wss.py).

Each of these checks fails on a plausible mistake planted in the C restatement: the TRT magic inverted, sigma on the
wrong term or the blend's sign flipped, the odd Guo factor with om+ instead of om-, the sign of the Galilean term or
of the third-order term, the sign of pref in the relaxation, the 1/3 of the probe pressure.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liblbm_oracle.so")
LIB_PATH_FMA = os.path.join(HERE, "_build", "liblbm_oracle_fma.so")
_libs = {}

_i32p = ctypes.POINTER(ctypes.c_int32)


def build():
    subprocess.run(["make", "-B", "-C", HERE], check=True, stdout=subprocess.DEVNULL)


def lib(fma=False):
    """The oracle's shared library (built on first use when absent). ``fma`` selects the build that contracts
    multiply-adds (as the reference's compiled FP32 kernels do); the default build keeps every rounding step."""
    if fma not in _libs:
        path = LIB_PATH_FMA if fma else LIB_PATH
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        for suf, rt in (("_f32", ctypes.c_float), ("_f64", ctypes.c_double)):
            rp = ctypes.POINTER(rt)
            getattr(L, "orc_clock" + suf).argtypes = [_i32p, rp, ctypes.c_int, ctypes.c_int, ctypes.c_int, rp]
            getattr(L, "orc_step" + suf).argtypes = [rp, rp, _i32p, rp, rp, _i32p, rp, rt, rp, rp, rt, ctypes.c_int,
                                                     ctypes.c_int, rt, ctypes.c_int, rp, rp, rp, ctypes.c_int64,
                                                     ctypes.c_int]
            getattr(L, "orc_run" + suf).argtypes = [rp, rp, _i32p, rp, rp, _i32p, rp, _i32p, rp, ctypes.c_int,
                                                    ctypes.c_int, ctypes.c_int, rp, rp, rp, rt, ctypes.c_int,
                                                    ctypes.c_int, rt, ctypes.c_int, rp, rp, rp, ctypes.c_int64,
                                                    ctypes.c_int, ctypes.c_int64, ctypes.c_int]
            getattr(L, "orc_run" + suf).restype = ctypes.c_int
            getattr(L, "orc_flux_bc" + suf).argtypes = [_i32p, rp, _i32p, ctypes.c_int64, rp, rp, rp, rt, rp,
                                                        ctypes.c_int, rt, rp, rp]
            getattr(L, "orc_gather" + suf).argtypes = [_i32p, _i32p, ctypes.c_int64, rp, rp, rp, rp]
        L.orc_tables.argtypes = [_i32p, ctypes.POINTER(ctypes.c_double), _i32p]
        _libs[fma] = L
    return _libs[fma]


def tables():
    """(C, W, OPP) as compiled into the oracle."""
    c = np.zeros((27, 3), np.int32)
    w = np.zeros(27, np.float64)
    o = np.zeros(27, np.int32)
    lib().orc_tables(_p(c), _p(w), _p(o))
    return c, w, o


def _p(a):
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError(f"oracle: array of shape {a.shape} and strides {a.strides} is not C-contiguous")
    return a.ctypes.data_as(ctypes.POINTER(np.ctypeslib.as_ctypes_type(a.dtype)))


def default_threads():
    return max(1, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count() or 1)


class OracleSolver:
    """CPU oracle with the reference Solver's state and stepping semantics (lbm.py:351-455).

    ``dtype`` float32 mirrors the reference's FP32 kernels; float64 is the FP64 error reference. State buffers are
    numpy arrays in the reference layout: ``f[2]`` (Q, N), ``vels[2]`` (N, 3), ``rho`` (N); ``cur`` selects the
    newest buffer. Graph chunking does not change results, so ``run`` simply steps.
    """

    def __init__(self, lat, nu_lu, magic=3.0 / 16.0, bouzidi=True, model="trt", sigma=1.0, galilean=None,
                 dtype=np.float32, threads=None, fma=False):
        self.fma = bool(fma)
        self.dtype = np.dtype(dtype)
        self.suf = "_f32" if self.dtype == np.float32 else "_f64"
        self.lat = lat
        self.N = N = lat["nbr"].shape[1]
        self.tau = 3.0 * nu_lu + 0.5
        self.magic = float(magic)
        self.bouzidi = int(bool(bouzidi))
        self.model = {"trt": 0, "regularised": 1}[model]
        self.sigma = float(sigma)
        self.gal = int(True if galilean is None else bool(galilean))
        self.threads = threads or default_threads()
        dt = self.dtype
        self.f = [np.zeros((27, N), dt), np.zeros((27, N), dt)]
        self.nbr = np.ascontiguousarray(lat["nbr"], np.int32)
        self.q = np.ascontiguousarray(lat["q"], dt)
        self.uin = np.ascontiguousarray(lat["uin"], dt)
        self.inl = np.ascontiguousarray((lat["nbr"] == -2).any(0), np.int32)
        self.p_out = np.zeros(max(1, lat["n_outlets"]), dt)
        self.vels = [np.zeros((N, 3), dt), np.zeros((N, 3), dt)]
        self.rho = np.zeros(N, dt)
        self.om = np.full(N, 1.0 / self.tau, dt)
        self.tick = np.zeros(1, np.int32)
        self.drv = np.zeros(1, dt)
        self.force = np.zeros(3, dt)
        self.set_drive([1.0])
        self.cur = 0
        self.steps = 0

    def set_drive(self, coef, period=1, ramp=0):
        coef = np.asarray(coef, np.float32)
        if len(coef) % 2 == 0:
            coef = np.append(coef, 0.0).astype(np.float32)
        self.coef = coef.astype(self.dtype)
        self.nh, self.period, self.ramp = (len(coef) - 1) // 2, int(period), int(ramp)

    def set_force(self, F):
        self.force = np.asarray([float(x) for x in F], np.float32).astype(self.dtype)

    def set_omega(self, om):
        self.om = np.ascontiguousarray(np.asarray(om, np.float32), self.dtype)

    def set_outlet_pressure(self, p_lu):
        self.p_out[:] = np.asarray(p_lu, np.float32).reshape(-1)

    def run(self, steps):
        L = lib(self.fma)
        fn = getattr(L, "orc_run" + self.suf)
        a = self.cur
        self.cur = fn(_p(self.f[0]), _p(self.f[1]), _p(self.nbr), _p(self.q), _p(self.uin), _p(self.inl),
                      _p(self.p_out), _p(self.tick), _p(self.coef), self.nh, self.period, self.ramp, _p(self.drv),
                      _p(self.force), _p(self.om), self.magic, self.bouzidi, self.model, self.sigma, self.gal,
                      _p(self.vels[0]), _p(self.vels[1]), _p(self.rho), self.N, a, int(steps), self.threads)
        self.steps += int(steps)

    def step_once(self, drive):
        """One k_step with an explicit drive value and no clock (for single-collision checks)."""
        fn = getattr(lib(self.fma), "orc_step" + self.suf)
        a = self.cur
        fn(_p(self.f[a]), _p(self.f[1 - a]), _p(self.nbr), _p(self.q), _p(self.uin), _p(self.inl), _p(self.p_out),
           self.dtype.type(drive), _p(self.force), _p(self.om), self.magic, self.bouzidi, self.model, self.sigma,
           self.gal, _p(self.vels[a]), _p(self.vels[1 - a]), _p(self.rho), self.N, self.threads)
        self.cur = 1 - a
        self.steps += 1

    def fields(self):
        return self.rho / 3.0, self.vels[self.cur]

    def get_f(self):
        return self.f[self.cur]

    def ok(self):
        return bool(np.isfinite(self.rho).all())


def clock(tick, coef, nh, period, ramp, dtype=np.float32):
    """One k_clock evaluation: returns (drive, new tick)."""
    dt = np.dtype(dtype)
    suf = "_f32" if dt == np.float32 else "_f64"
    t = np.array([tick], np.int32)
    c = np.ascontiguousarray(coef, dt)
    d = np.zeros(1, dt)
    getattr(lib(), "orc_clock" + suf)(_p(t), _p(c), int(nh), int(period), int(ramp), _p(d))
    return d[0], int(t[0])


def flux_bc(idx, w, slot, vel, Qacc, R, pref, drive, n_open, alpha, p_out, q_last, dtype=np.float32):
    dt = np.dtype(dtype)
    suf = "_f32" if dt == np.float32 else "_f64"
    getattr(lib(), "orc_flux_bc" + suf)(_p(idx), _p(w), _p(slot), len(idx), _p(vel), _p(Qacc), _p(R),
                                        dt.type(pref), _p(drive), int(n_open), dt.type(alpha), _p(p_out), _p(q_last))


def gather(idx, slot, rho, vel, out, umax, dtype=np.float32):
    dt = np.dtype(dtype)
    suf = "_f32" if dt == np.float32 else "_f64"
    getattr(lib(), "orc_gather" + suf)(_p(idx), _p(slot), len(idx), _p(rho), _p(vel), _p(out), _p(umax))


def sponge_tau(lat, tau0, length, tau_max, kinds=("inlet", "outlet")):
    """Per-cell tau of the sponge layers (restates lbm.py:398-416): tau_max at the opening plane, tau0 at `length`
    cells inside, smoothstep between, within 1.5 radii of the opening axis."""
    P = lat["origin"] + lat["coords"] * lat["dx"]
    tau = np.full(lat["nbr"].shape[1], tau0)
    for o in lat["openings"]:
        if o["kind"] not in kinds:
            continue
        c, nrm = np.asarray(o["centre"], float), np.asarray(o["normal"], float)
        rel = P - c
        depth = -(rel @ nrm) / lat["dx"]
        side = np.linalg.norm(rel - (rel @ nrm)[:, None] * nrm, axis=1) < 1.5 * o["radius"]
        w = np.clip(1.0 - depth / length, 0.0, 1.0)
        w = w * w * (3.0 - 2.0 * w)
        tau = np.maximum(tau, np.where(side, tau0 + (tau_max - tau0) * w, tau0))
    return tau


def _oracle_set_sponge(self, length, tau_max, kinds=("inlet", "outlet")):
    self.tau_cell = sponge_tau(self.lat, self.tau, length, tau_max, kinds)
    self.set_omega((1.0 / self.tau_cell).astype(np.float32))


OracleSolver.set_sponge = _oracle_set_sponge
