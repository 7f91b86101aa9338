"""Sparse D3Q27 lattice-Boltzmann solver on MI355X -- drop-in for CerebroSim's ``lbm`` module (lbm.py).

Same module surface as the reference: ``LATTICE, C, W, Q, OPP, WALL, INLET, OUTLET0, Units, build_lattice, scatter,
Solver`` (+ ``synchronize``, the counterpart of ``wp.synchronize`` that the reference's callers use). The solver
step runs as hand-written HIP kernels for gfx950 through the C ABI of libcerebro_lbm.so; PyTorch-ROCm tensors are
only device buffers and the stream. There is no CPU execution path: constructing a Solver without a GPU or without
the built library raises.

Solver semantics follow lbm.py:346-455: per step the drive clock then the step kernel; ``run(steps, chunk)`` replays
a captured graph of ``chunk`` (made even) steps and launches the remainder one by one; ``set_drive``/``set_force``
invalidate the graphs; ``set_sponge``/``set_outlet_pressure`` update device buffers in place. ``_launch(a)`` is the
per-step hook that subclasses extend (simulate.py:109-114 appends its flow/boundary kernels there); it is captured
into the graphs like the reference's.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native
from .lattice import C, INLET, LATTICE, OPP, OUTLET0, Q, WALL, W, Units, build_lattice, scatter  # noqa: F401

__all__ = ["LATTICE", "C", "W", "Q", "OPP", "WALL", "INLET", "OUTLET0", "Units", "build_lattice", "scatter",
           "Solver", "synchronize"]


def _torch():
    import torch
    return torch


def synchronize():
    """Wait for all device work (counterpart of wp.synchronize)."""
    _native.call("cs_device_synchronize")


def _require_gpu(device):
    torch = _torch()
    _native.load()
    if not torch.cuda.is_available():
        raise RuntimeError("cerebrosim_amd.lbm.Solver needs a ROCm GPU (torch.cuda.is_available() is False); "
                           "there is no CPU execution path")
    dev = torch.device(device if device is not None else "cuda:0")
    if dev.type != "cuda":
        raise RuntimeError(f"device {device!r}: the solver runs on the GPU only")
    return dev


class _IndexedBackend:
    """Reference layout on the device: f[2] (Q, N), nbr/q (Q, N), uin/vels (N, 3) -- any neighbour table."""

    name = "indexed"

    def __init__(self, s, lat):
        torch = _torch()
        dev, N = s.dev, s.N
        self.f = [torch.zeros((Q, N), dtype=torch.float32, device=dev) for _ in range(2)]
        self.nbr = torch.from_numpy(np.ascontiguousarray(lat["nbr"], np.int32)).to(dev)
        self.q = torch.from_numpy(np.ascontiguousarray(lat["q"], np.float32)).to(dev)
        self.uin = torch.from_numpy(np.ascontiguousarray(lat["uin"], np.float32)).to(dev)
        self.inl = torch.from_numpy((lat["nbr"] == INLET).any(0).astype(np.int32)).to(dev)
        self.vels = [torch.zeros((N, 3), dtype=torch.float32, device=dev) for _ in range(2)]
        self.rho = torch.zeros(N, dtype=torch.float32, device=dev)
        self.om = torch.full((N,), 1.0 / s.tau, dtype=torch.float32, device=dev)

    def set_omega(self, om):
        self.om.copy_(_torch().from_numpy(np.ascontiguousarray(om, np.float32)))

    def density(self, cur):
        return self.rho

    def density_stride(self):
        return 1

    def finite_view(self, cur):
        return self.rho

    def fields(self, cur):
        return self.rho.cpu().numpy(), self.vels[cur].cpu().numpy()

    def launch_parts(self):
        return [("k_step_indexed", self.vels[0].shape[0])]

    def step_part(self, s, a, part):
        self.step(s, a)

    def step(self, s, a):
        args = _native.StepArgs()
        args.f_in, args.f_out = self.f[a].data_ptr(), self.f[1 - a].data_ptr()
        args.nbr, args.qw, args.uin, args.inl = (self.nbr.data_ptr(), self.q.data_ptr(), self.uin.data_ptr(),
                                                 self.inl.data_ptr())
        args.p_out, args.drive, args.om_cell = s.p_out.data_ptr(), s._drive_ptr(), self.om.data_ptr()
        args.vel_old, args.vel, args.rho = self.vels[a].data_ptr(), self.vels[1 - a].data_ptr(), self.rho.data_ptr()
        args.n_cells = s.N
        args.force = (ctypes.c_float * 3)(*s.force)
        args.magic, args.sigma = s.magic, s.sigma
        args.bouzidi, args.model, args.galilean = s.bouzidi, s.model, s.gal
        _native.call("cs_step_indexed", ctypes.byref(args), s.stream_handle)

    # state access in the reference layout
    def get_f(self, s, cur):
        return self.f[cur].cpu().numpy()

    def set_f(self, s, cur, f):
        self.f[cur].copy_(_torch().from_numpy(np.ascontiguousarray(f, np.float32)))

    def velocity(self, cur):
        return self.vels[cur]

    def velocity_addressing(self):
        return 3, 1            # (element stride, component stride) in floats

    def slots(self, cells):
        return np.asarray(cells, np.int32)

    def set_velocity(self, cur, u, s=None):
        self.vels[cur].copy_(_torch().from_numpy(np.ascontiguousarray(u, np.float32)))


class BrickMoments:
    """Per-slot moments of a brick storage (S slots), double-buffered by step parity.

    FP32 storage: velocity as three planes (3, S) per parity and one density plane (the reference's separate vel /
    rho arrays). FP16 storage: one float4 {drho, ux, uy, uz} per slot (S, 4) per parity -- the decoding of a
    population needs the writer's previous density and velocity together, so one 16-byte load fetches them."""

    def __init__(self, S, fp16, dev):
        torch = _torch()
        self.S, self.fp16 = int(S), bool(fp16)
        if self.fp16:
            self.m = [torch.zeros((self.S, 4), dtype=torch.float32, device=dev) for _ in range(2)]
        else:
            self.v = [torch.zeros((3, self.S), dtype=torch.float32, device=dev) for _ in range(2)]
            self.r = torch.zeros(self.S, dtype=torch.float32, device=dev)

    def tensors(self):
        return list(self.m) if self.fp16 else list(self.v) + [self.r]

    def step_ptrs(self, a):
        """(vel_old, vel, rho) of a step of parity a."""
        if self.fp16:
            return self.m[a].data_ptr(), self.m[1 - a].data_ptr(), 0
        return self.v[a].data_ptr(), self.v[1 - a].data_ptr(), self.r.data_ptr()

    def convert_ptrs(self, cur):
        """(vel, rho) of the newest moments, for the conversions."""
        if self.fp16:
            return self.m[cur].data_ptr(), 0
        return self.v[cur].data_ptr(), self.r.data_ptr()

    def velocity(self, cur):
        return self.m[cur][:, 1:] if self.fp16 else self.v[cur]

    def velocity_addressing(self):
        """(element stride, component stride) in floats from velocity(cur).data_ptr()."""
        return (4, 1) if self.fp16 else (1, self.S)

    def density(self, cur):
        return self.m[cur][:, 0] if self.fp16 else self.r

    def density_stride(self):
        return 4 if self.fp16 else 1

    def flat(self, cur, lo=0, hi=None):
        """Contiguous float view of the newest moments of slots [lo, hi) whose finiteness ok() checks."""
        hi = self.S if hi is None else hi
        return self.m[cur][lo:hi].reshape(-1) if self.fp16 else self.r[lo:hi]

    def gather(self, cur, idx):
        """(drho (n,), u (n, 3)) at slots idx (device int64)."""
        if self.fp16:
            m = self.m[cur].index_select(0, idx)
            return m[:, 0].contiguous(), m[:, 1:].contiguous()
        return self.r.index_select(0, idx), self.v[cur].index_select(1, idx).t().contiguous()

    def set_velocity(self, cur, idx, u):
        if self.fp16:
            self.m[cur][idx, 1:] = u
        else:
            self.v[cur][:, idx] = u.t()

    def zero(self, lo, hi):
        for t in self.tensors():
            if t.dim() == 2 and not self.fp16:
                t[:, lo:hi].zero_()
            else:
                t[lo:hi].zero_()


class _BricksBackend:
    """4^3 brick storage with AA in-place streaming (see csrc/k_bricks.hip): populations [bricks][27][64] once,
    moments per slot (BrickMoments). Two launches per step: the interior bricks (no boundary data) and the boundary
    bricks.

    storage "fp32": populations stored as float. "fp16": stored as float16 deviations from the writing cell's
    second-order equilibrium (scaled by 1024 / w_i); the arithmetic stays float32, the side buffer float32."""

    name = "bricks"

    def __init__(self, s, lat, layout, storage="fp32"):
        torch = _torch()
        dev = s.dev
        self.L = layout
        self.S = S = int(layout["n_slots"])
        self.nb = int(layout["n_bricks"])
        self.n_cells = s.N
        self.storage = {"fp32": _native.CS_STORAGE_FP32, "fp16": _native.CS_STORAGE_FP16}[storage]
        up = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
        dt = torch.float16 if self.storage == _native.CS_STORAGE_FP16 else torch.float32
        self.f = torch.zeros(self.nb * 27 * 64, dtype=dt, device=dev)
        self.nbr = up(layout["nbr"])
        self.bidx = up(layout["bidx"])
        self.interior = up(layout["interior"])
        self.boundary = up(layout["boundary"])
        self.lmask = up(layout["lmask"].view(np.int32)) if layout["n_boundary"] else torch.zeros(64, dtype=torch.int32, device=dev)
        self.smask = up(layout["smask"].view(np.int32)) if layout["n_boundary"] else torch.zeros(64, dtype=torch.int32, device=dev)
        self.qlink = up(layout["qlink"])                        # padded by one element (kernels read ahead)
        self.qstart = up(layout["qstart"])
        self.sstart = up(layout["sstart"])
        nbd = int(layout["n_boundary"])
        self.loff = up(layout["loff"].view(np.int32)) if nbd else torch.zeros(64, dtype=torch.int32, device=dev)
        self.ldir = up(layout["ldir"].view(np.int32)) if nbd else torch.zeros(2, dtype=torch.int32, device=dev)
        self.uin = up(layout["uin"]) if layout["n_boundary"] else torch.zeros(3 * 64, dtype=torch.float32, device=dev)
        self.side = torch.zeros(layout["n_side"] + 1, dtype=torch.float32, device=dev)
        self.slot_of_cell = up(layout["slot_of_cell"].astype(np.int64))
        self.slot_of_cell32 = up(layout["slot_of_cell"])
        self.mom = BrickMoments(S, self.storage == _native.CS_STORAGE_FP16, dev)
        self.omega = float(np.float32(1.0 / s.tau))
        self.om = torch.full((S,), self.omega, dtype=torch.float32, device=dev)
        self.om_brick = torch.zeros(self.nb, dtype=torch.uint8, device=dev)

    def set_omega(self, om):
        torch = _torch()
        om = np.asarray(om, np.float32)
        slots = np.full(self.S, self.omega, np.float32)
        slots[self.L["slot_of_cell"]] = om
        own = (slots.reshape(self.nb, 64) != np.float32(self.omega)).any(1).astype(np.uint8)
        self.om.copy_(torch.from_numpy(slots))
        self.om_brick.copy_(torch.from_numpy(own))

    def _args(self, s, a, boundary):
        A = _native.BricksStepArgs()
        lst = self.boundary if boundary else self.interior
        A.f, A.nbr, A.list, A.n_list = self.f.data_ptr(), self.nbr.data_ptr(), lst.data_ptr(), lst.numel()
        A.bidx, A.lmask, A.smask = self.bidx.data_ptr(), self.lmask.data_ptr(), self.smask.data_ptr()
        A.qlink, A.qstart, A.uin = self.qlink.data_ptr(), self.qstart.data_ptr(), self.uin.data_ptr()
        A.side, A.sstart = self.side.data_ptr(), self.sstart.data_ptr()
        A.loff, A.ldir = self.loff.data_ptr(), self.ldir.data_ptr()
        A.om_cell, A.om_brick, A.omega = self.om.data_ptr(), self.om_brick.data_ptr(), self.omega
        A.p_out, A.drive = s.p_out.data_ptr(), s._drive_ptr()
        A.vel_old, A.vel, A.rho = self.mom.step_ptrs(a)
        A.n_slots, A.storage = self.S, self.storage
        A.force = (ctypes.c_float * 3)(*s.force)
        A.magic = s.magic
        A.bouzidi, A.model, A.galilean = s.bouzidi, s.model, s.gal
        A.hybrid, A.sigma = int(s.model == 1 and s.sigma < 1.0), s.sigma
        A.parity, A.boundary = a, int(boundary)
        A.variant = s.variant
        return A

    def step(self, s, a):
        """The step's bricks in one launch (cs_bricks_step_pair): the boundary and interior lists are independent
        within a step (in the AA pattern every slot is read and then written by one thread of the step, interior or
        boundary brick alike; the moments read are the previous step's; the Bouzidi side copies are a boundary
        cell's own), so one grid holds both, the boundary workgroups first (the boundary kernels are bound by
        instruction issue and latency, the interior ones by HBM)."""
        if not (self.L["n_interior"] and self.L["n_boundary"]):
            for boundary in (False, True):
                self.step_part(s, a, boundary)
            return
        Ab, Ai = self._args(s, a, True), self._args(s, a, False)
        _native.call("cs_bricks_step_pair", ctypes.byref(Ab), ctypes.byref(Ai), s.stream_handle)

    def step_part(self, s, a, boundary, stream=None):
        """One of the step's two launches: interior bricks or boundary bricks (on the solver stream by default)."""
        A = self._args(s, a, boundary)
        if A.n_list:
            _native.call("cs_bricks_step", ctypes.byref(A), stream if stream is not None else s.stream_handle)

    def launch_parts(self):
        """(name, cells processed) of the launches of one step, in order."""
        return [("k_bricks_interior", int(self.L["n_interior"]) * 64),
                ("k_bricks_boundary", self.n_cells - int(self.L["n_interior"]) * 64)]

    def _convert(self, s, name, f_ref, parity):
        A = _native.BricksConvertArgs()
        A.f, A.nbr, A.bidx = self.f.data_ptr(), self.nbr.data_ptr(), self.bidx.data_ptr()
        A.lmask, A.smask, A.side = self.lmask.data_ptr(), self.smask.data_ptr(), self.side.data_ptr()
        A.sstart, A.loff = self.sstart.data_ptr(), self.loff.data_ptr()
        A.slot_of_cell, A.n_cells = self.slot_of_cell32.data_ptr(), self.n_cells
        A.f_ref, A.parity = f_ref.data_ptr(), parity
        # the newest moments: those of the step that left this state (parity = the next step's parity)
        A.vel, A.rho = self.mom.convert_ptrs(parity)
        A.storage, A.n_slots = self.storage, self.S
        _native.call(name, ctypes.byref(A), s.stream_handle)

    def get_f(self, s, cur):
        f_ref = _torch().empty((Q, self.n_cells), dtype=_torch().float32, device=s.dev)
        self._convert(s, "cs_bricks_get_f", f_ref, cur)
        s.synchronize()
        return f_ref.cpu().numpy()

    def set_f(self, s, cur, f):
        f_ref = _torch().from_numpy(np.ascontiguousarray(f, np.float32)).to(s.dev)
        self._convert(s, "cs_bricks_set_f", f_ref, cur)
        s.synchronize()

    def velocity(self, cur):
        return self.mom.velocity(cur)

    def velocity_addressing(self):
        return self.mom.velocity_addressing()

    def slots(self, cells):
        return self.L["slot_of_cell"][np.asarray(cells)].astype(np.int32)

    def density(self, cur):
        return self.mom.density(cur)

    def density_stride(self):
        return self.mom.density_stride()

    def finite_view(self, cur):
        return self.mom.flat(cur)

    def fields(self, cur):
        rho, u = self.mom.gather(cur, self.slot_of_cell)
        return rho.cpu().numpy(), u.cpu().numpy()

    def set_velocity(self, cur, u, s=None):
        torch = _torch()
        # FP16: the stored deviations are relative to these moments -- re-encode the populations around the new ones
        f = self.get_f(s, cur) if self.storage == _native.CS_STORAGE_FP16 else None
        u = torch.from_numpy(np.ascontiguousarray(u, np.float32)).to(self.f.device)
        self.mom.set_velocity(cur, self.slot_of_cell, u)
        if f is not None:
            self.set_f(s, cur, f)


def sponge_tau(lat, tau0, length, tau_max, kinds=("inlet", "outlet")):
    """Per-cell tau with sponge layers at the openings of the given kinds (lbm.py:398-416). The side test and the
    smoothstep run only on the cells closer to the opening's plane than ``length`` (elsewhere the weight is 0 and tau
    keeps its value); each cell's arithmetic is the full formula's."""
    P = lat["origin"] + lat["coords"] * lat["dx"]
    tau = np.full(lat["nbr"].shape[1], tau0)
    for o in lat["openings"]:
        if o["kind"] not in kinds:
            continue
        centre = np.asarray(o["centre"], float)
        normal = np.asarray(o["normal"], float)
        rel = P - centre
        along = rel @ normal
        depth = -along / lat["dx"]
        near = np.nonzero(depth < length)[0] if length > 0 else np.arange(len(depth))
        if not near.size:
            continue
        r, a, d = rel[near], along[near], depth[near]
        side = np.linalg.norm(r - a[:, None] * normal, axis=1) < 1.5 * o["radius"]
        w = np.clip(1.0 - d / length, 0.0, 1.0)
        w = w * w * (3.0 - 2.0 * w)
        tau[near] = np.maximum(tau[near], np.where(side, tau0 + (tau_max - tau0) * w, tau0))
    return tau


class Solver:
    """Lattice Boltzmann on a sparse lattice (lbm.py:346-455). ``model`` "trt" (magic = Lambda) or "regularised";
    ``sigma`` < 1 blends the regularised stress with the finite-difference one. ``layout`` selects the device
    storage: "bricks" (4^3 bricks, AA in-place streaming), "indexed" (reference (Q, N) layout, any neighbour table)
    or "auto" (bricks when the lattice and options allow, else indexed). ``storage`` "fp16" keeps the populations
    in float16 (bricks layout only, not with the hybrid blend; float32 arithmetic and reductions).
    ``brick_layout``: the host layout of ``bricks.build(lat)`` when the caller already has it."""

    def __init__(self, lat, nu_lu, magic=3.0 / 16.0, bouzidi=True, model="trt", sigma=1.0, galilean=None,
                 device=None, layout="auto", storage="fp32", brick_layout=None):
        torch = _torch()
        self.dev = _require_gpu(device)
        self.lat = lat
        self.N = N = lat["nbr"].shape[1]
        self.tau = 3.0 * nu_lu + 0.5
        self.magic = float(magic)
        self.bouzidi = int(bool(bouzidi))
        self.model = {"trt": 0, "regularised": 1}[model]
        self.sigma = float(sigma)
        self.gal = int(LATTICE == "D3Q27" if galilean is None else bool(galilean))
        if self.gal and LATTICE != "D3Q27":
            raise ValueError("the Galilean correction needs D3Q27")
        if LATTICE != "D3Q27":
            raise NotImplementedError("the HIP kernels implement D3Q27 only")
        dev = self.dev
        self._tstream = torch.cuda.Stream(device=dev)          # default priority
        self.stream_handle = ctypes.c_void_p(self._tstream.cuda_stream)
        with torch.cuda.stream(self._tstream):
            self.p_out = torch.zeros(max(1, lat["n_outlets"]), dtype=torch.float32, device=dev)
            self.tick = torch.zeros(1, dtype=torch.int32, device=dev)
            self._clk = torch.zeros(3, dtype=torch.int32, device=dev)          # nh, period, ramp
            self._dtab = torch.zeros(256, dtype=torch.float32, device=dev)     # drives of the steps of a chunk
            self._k = 0
            self._work = torch.zeros(_native.CS_REDUCE_WORK, dtype=torch.float64, device=dev)
            self._red = torch.zeros(2, dtype=torch.float64, device=dev)
            if storage not in ("fp32", "fp16"):
                raise ValueError(f"unknown storage {storage!r}")
            self.storage = storage
            self.backend = self._make_backend(lat, layout, brick_layout)
        self._tstream.synchronize()
        self.layout = self.backend.name
        self.force = (0.0, 0.0, 0.0)
        self.variant = 0                   # launch variant of the brick kernels (cs_bricks_variant_info)
        self.graphs = {}
        self.set_drive([1.0])
        self.cur = 0
        self.steps = 0

    def _make_backend(self, lat, layout, brick_layout=None):
        from . import bricks
        if layout not in ("auto", "bricks", "indexed"):
            raise ValueError(f"unknown layout {layout!r}")
        if layout == "indexed":
            if self.storage != "fp32":
                raise ValueError("storage 'fp16' needs the bricks layout")
            return _IndexedBackend(self, lat)
        reason = None
        if self.model == 1 and self.sigma < 1.0 and self.storage != "fp32":
            reason = "the hybrid blend (sigma < 1) runs in FP32 storage"
        else:
            try:
                lay = brick_layout if brick_layout is not None else bricks.build(lat)
                return _BricksBackend(self, lat, lay, self.storage)
            except bricks.LayoutError as e:
                reason = str(e)
        if layout == "bricks" or self.storage != "fp32":
            raise ValueError(f"layout 'bricks' (needed for storage {self.storage!r}) unavailable: {reason}")
        self.layout_note = reason
        return _IndexedBackend(self, lat)

    @property
    def om(self):
        return self.backend.om

    # --- reference attributes backed by the device state
    @property
    def vels(self):
        return [self.backend.velocity(0), self.backend.velocity(1)]

    @property
    def rho(self):
        return self.backend.density(self.cur)

    @property
    def f(self):
        return self.backend.f

    # --- inputs (lbm.py:384-419)
    def _host_to_device(self, dst, array):
        torch = _torch()
        self._tstream.synchronize()
        with torch.cuda.stream(self._tstream):
            dst.copy_(torch.from_numpy(np.ascontiguousarray(array)).to(self.dev, non_blocking=False))
        self._tstream.synchronize()

    def set_drive(self, coef, period=1, ramp=0):
        """drive(t) = coef[0] + sum_k coef[2k-1] cos(2 pi k t / period) + coef[2k] sin(...) scales the inlet profile
        and the body force; ramp: smooth start over that many steps. Invalidates the captured graphs."""
        torch = _torch()
        coef = np.asarray(coef, np.float32)
        if len(coef) % 2 == 0:
            coef = np.append(coef, 0.0).astype(np.float32)
        self._invalidate()
        self.coef = torch.from_numpy(coef).to(self.dev)
        self.nh, self.period, self.ramp = (len(coef) - 1) // 2, int(period), int(ramp)
        with torch.cuda.stream(self._tstream):
            self._clk.copy_(torch.tensor([self.nh, self.period, self.ramp], dtype=torch.int32))
        self._tstream.synchronize()

    def set_force(self, F):
        self.force = tuple(float(np.float32(x)) for x in F)
        self._invalidate()

    def _invalidate(self):
        """Drop the captured graphs (after the stream drained, so none is still executing)."""
        self._tstream.synchronize()
        self.graphs = {}

    def set_sponge(self, length, tau_max, kinds=("inlet", "outlet")):
        """Raise tau smoothly towards the openings of the given kinds: tau_max at the opening plane, back to the
        fluid's tau `length` cells inside, within 1.5 radii of the opening axis (lbm.py:398-416). Written in place."""
        tau = sponge_tau(self.lat, self.tau, length, tau_max, kinds)
        self.tau_cell = tau
        self.synchronize()
        with _torch().cuda.stream(self._tstream):
            self.backend.set_omega((1.0 / tau).astype(np.float32))
        self.synchronize()

    def set_outlet_pressure(self, p_lu):
        self._host_to_device(self.p_out, np.asarray(p_lu, np.float32).reshape(-1))

    # --- stepping (lbm.py:422-448)
    @property
    def drv(self):
        """Drive of the most recent step (1-element device tensor; lbm.py:379)."""
        return self._dtab[self._k:self._k + 1]

    def _drive_ptr(self):
        """Device address of the drive the step being launched reads."""
        return self._dtab.data_ptr() + 4 * self._k

    def _clock_table(self, n):
        """k_clock for the next n steps at once (lbm.py:90-102): drives into the table, tick += n."""
        if n > self._dtab.numel():
            torch = _torch()
            self._invalidate()
            with torch.cuda.stream(self._tstream):
                self._dtab = torch.zeros(n, dtype=torch.float32, device=self.dev)
        A = _native.ClockTableArgs()
        A.tick, A.coef, A.drive = self.tick.data_ptr(), self.coef.data_ptr(), self._dtab.data_ptr()
        base = self._clk.data_ptr()
        A.nh, A.period, A.ramp = base, base + 4, base + 8
        A.n_cases, A.coef_stride, A.n_steps, A.drive_stride = 1, self.coef.numel(), int(n), 1
        _native.call("cs_clock_table", ctypes.byref(A), self.stream_handle)

    def _clock(self):
        """The clock of one step (eager launches)."""
        self._clock_table(1)
        self._k = 0

    def _launch(self, a):
        """The launches of one step of parity a, reading the drive at table slot self._k (the clock runs once per
        chunk, before the chunk's steps)."""
        self.backend.step(self, a)

    def run(self, steps, chunk=200):
        """Advance `steps` steps: graphs of `chunk` steps (made even), remainder launched step by step."""
        chunk += chunk % 2
        full, rest = divmod(int(steps), chunk)
        if full:
            g = self.prepare(chunk)
            for _ in range(full):
                g.launch(self.stream_handle)
            self._k = chunk - 1
        for _ in range(rest):
            self._clock()
            self._launch(self.cur)
            self.cur = 1 - self.cur
        self.steps += int(steps)

    def prepare(self, chunk):
        """The graph of `chunk` (even) steps from the current parity, captured now if it is not cached (capture
        launches nothing); run() replays it. Returns the graph."""
        chunk = int(chunk)
        if chunk <= 0 or chunk % 2:
            raise ValueError(f"graph chunk {chunk}: must be even and positive")
        if chunk > self._dtab.numel():          # grow the table outside the capture
            self._invalidate()
            torch = _torch()
            with torch.cuda.stream(self._tstream):
                self._dtab = torch.zeros(chunk, dtype=torch.float32, device=self.dev)
        key = (self.cur, chunk)
        if key not in self.graphs:
            with _native.capture(self.stream_handle) as cap:
                self._clock_table(chunk)
                a = self.cur
                for k in range(chunk):
                    self._k = k
                    self._launch(a)
                    a = 1 - a
            self.graphs[key] = cap.graph
        return self.graphs[key]

    def synchronize(self):
        self._tstream.synchronize()

    def autotune(self, steps=20, repeats=3):
        """Time each launch variant of the brick kernels (workgroup size, XCD order) on this lattice and keep the
        fastest; the timing steps are undone (state restored). Each variant is timed as run() steps it: replays of a
        captured graph of `steps` steps (made even; one untimed replay first, then the best of `repeats`). Returns
        {variant: ms per step}. All variants give the same results."""
        if self.layout != "bricks":
            return {}
        from .batch import launch_variants
        torch = _torch()
        b = self.backend
        # the snapshot and the restore run on the solver stream, after the work already queued there
        state = [b.f, b.side] + b.mom.tensors() + [self.tick, self.p_out] + self._stepped_state()
        self.synchronize()
        with torch.cuda.stream(self._tstream):
            saved = [t.clone() for t in state]
        cur, nsteps = self.cur, self.steps
        steps = max(2, int(steps) + int(steps) % 2)
        times = {}
        for v in range(len(launch_variants())):
            self._invalidate()
            self.variant = v
            g = self.prepare(steps)
            g.launch(self.stream_handle)
            best = None
            for _ in range(repeats):
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(self._tstream)
                g.launch(self.stream_handle)
                e1.record(self._tstream)
                self.synchronize()
                ms = e0.elapsed_time(e1) / steps
                best = ms if best is None else min(best, ms)
            times[v] = best
            del g
        self._invalidate()
        with torch.cuda.stream(self._tstream):
            for dst, src in zip(state, saved):
                dst.copy_(src)
        self.synchronize()
        del saved
        self.cur, self.steps = cur, nsteps
        self.variant = min(times, key=times.get)
        self.graphs = {}
        return times

    def _stepped_state(self):
        """Device state beyond the populations, moments, clock and opening pressures that steps change (subclasses
        add theirs); autotune restores it."""
        return []

    # --- outputs (lbm.py:450-455)
    def fields(self):
        """(pressure, velocity) per cell in lattice units, as numpy, in the lattice's cell order."""
        self.synchronize()
        with _torch().cuda.stream(self._tstream):
            p, u = self.backend.fields(self.cur)
        return p / 3.0, u

    def ok(self):
        """All densities finite (device reduction; reads back 16 bytes)."""
        x = self.backend.finite_view(self.cur)
        _native.call("cs_reduce_finite", ctypes.c_void_p(x.data_ptr()), x.numel(),
                     ctypes.c_void_p(self._work.data_ptr()), ctypes.c_void_p(self._red.data_ptr()),
                     self.stream_handle)
        self.synchronize()
        return bool(self._red[1].item() == 0)

    # --- state access beyond the reference API (parity tests, fault injection, snapshots)
    def get_f(self):
        """Newest post-collision populations f~ (Q, N) in the reference layout, as numpy."""
        self.synchronize()
        with _torch().cuda.stream(self._tstream):
            return self.backend.get_f(self, self.cur)

    def set_f(self, f):
        """Replace the newest post-collision populations (Q, N), reference layout."""
        self.synchronize()
        with _torch().cuda.stream(self._tstream):
            self.backend.set_f(self, self.cur, f)
        self.synchronize()

    def set_velocity(self, u):
        """Replace the newest velocity field (N, 3) (the previous-step velocity of the next step)."""
        self.synchronize()
        with _torch().cuda.stream(self._tstream):
            self.backend.set_velocity(self.cur, u, self)
        self.synchronize()
