"""Batching form: many lattices (cases) of different sizes stepped together by one launch per brick arena.

Design (MI355X-first):

* An ``Arena`` is one packed brick storage on the device: populations ``[bricks][1728]`` (float or half), neighbour
  table, velocity / density planes, per-cell omega, and the boundary records. Cases take a contiguous range of
  bricks and a contiguous range of boundary records from first-fit allocators; their neighbour ids are offset by
  the range start on upload, so the kernels address the arena exactly as they address one lattice.
* The brick lists of an arena (interior, boundary) hold ``brick | case << 22``; the kernel reads the case's drive,
  base omega, body force and opening pressures from per-case tables. Both lists hold live entries only; beside the
  boundary list, ``brec`` gives each entry's boundary record (records keep their places while cases come and go).
* One step = one clock launch for all cases + two step launches per arena (interior, boundary; the boundary ones on
  a side stream, concurrent with the interior ones) + one flow launch per arena when cases measure flows
  (simulate.py's k_flux + k_bc). ``run`` captures chunks of steps in a graph. The launches cover the live extent
  of each list, so a graph is captured again when cases join or leave (a capture costs about as much as launching
  the chunk's steps one by one).
* An arena holds at most 2^32 / 1728 - 1 bricks (32-bit population offsets, about 159 M cells); a batch holds
  several arenas (one storage type each). Cases are independent, so results do not depend on what else is in the
  batch: the tests check that a case gives the same bits alone and in a mix.

The case handle (``Case``) has the ``lbm.Solver`` surface used by the callers: ``set_drive``, ``set_force``,
``set_sponge``, ``set_outlet_pressure``, ``fields``, ``get_f``/``set_f``, ``set_velocity``, ``ok``, plus the
resistance-outlet flow measurement of ``artery.ArterySolver`` (``set_outlets``, ``flows``, ``opening_pressures``).
"""
from __future__ import annotations

import collections
import ctypes
import time

import numpy as np

from . import _native
from .lbm import BrickMoments, _require_gpu, sponge_tau

BC = 64
BLOCK = 27 * BC
MAX_BRICKS = (2 ** 32 - 1) // BLOCK          # 32-bit population offsets per arena
STORAGES = {"fp32": _native.CS_STORAGE_FP32, "fp16": _native.CS_STORAGE_FP16}
MODELS = {"trt": 0, "regularised": 1}


def _torch():
    import torch
    return torch


def launch_variants():
    """The launch variants of the brick kernels: [{"variant", "waves_per_group", "xcd_order"}]."""
    out = []
    v = 0
    while True:
        w, x = ctypes.c_int32(), ctypes.c_int32()
        if _native.load().cs_bricks_variant_info(v, ctypes.byref(w), ctypes.byref(x)) != 0:
            break
        out.append({"variant": v, "waves_per_group": w.value, "xcd_order": x.value})
        v += 1
    return out


class BatchFull(RuntimeError):
    """No arena has room for the case (bricks, boundary records or case slots)."""


class _Ranges:
    """First-fit allocator of contiguous ranges in [0, capacity)."""

    def __init__(self, capacity):
        self.capacity = int(capacity)
        self.free = [(0, self.capacity)] if capacity > 0 else []

    def alloc(self, n):
        if n == 0:
            return 0
        for k, (s, ln) in enumerate(self.free):
            if ln >= n:
                if ln == n:
                    self.free.pop(k)
                else:
                    self.free[k] = (s + n, ln - n)
                return s
        return None

    def release(self, start, n):
        if n == 0:
            return
        self.free.append((start, n))
        self.free.sort()
        merged = []
        for s, ln in self.free:
            if merged and merged[-1][0] + merged[-1][1] == s:
                merged[-1] = (merged[-1][0], merged[-1][1] + ln)
            else:
                merged.append((s, ln))
        self.free = merged

    def high(self):
        """End of the highest allocated range."""
        if not self.free or self.free[-1][0] + self.free[-1][1] != self.capacity:
            return self.capacity
        return self.free[-1][0]


LINKS_PER_RECORD = 320        # default capacity of the compact link pool per boundary record (a sphere-like
SIDES_PER_RECORD = 160        # wall has ~125 missing links and ~60 side copies per boundary brick)


def arena_bytes(bricks, records, storage, links=None, sides=None):
    """Device bytes of an arena with the given capacities (link / side pools default to the per-record figures)."""
    links = LINKS_PER_RECORD * records if links is None else links
    sides = SIDES_PER_RECORD * records if sides is None else sides
    elem = 2 if storage == "fp16" else 4
    per_brick = BLOCK * elem + 27 * 4 + 4 + 1 + BC * 4 + 2 * 3 * BC * 4 + BC * 4 * (2 if storage == "fp16" else 1) + 4
    per_record = BC * (4 + 4 + 3 * 4 + 4) + 2 * 4 + 2 * 4
    return int(bricks) * per_brick + int(records) * per_record + 4 * (int(links) + int(sides) + 2)


def case_bytes(layout, storage):
    """Device bytes a case occupies inside an arena (its bricks, boundary records and link / side entries)."""
    return arena_bytes(layout["n_bricks"], layout["n_boundary"], storage, layout["n_links"], layout["n_side"])


class Arena:
    """Packed brick storage for the cases of one storage type."""

    def __init__(self, batch, bricks, records, links, sides):
        torch = _torch()
        if not 0 < bricks <= MAX_BRICKS:
            raise ValueError(f"arena of {bricks} bricks: 1 .. {MAX_BRICKS} (32-bit population offsets)")
        dev = batch.dev
        self.batch = batch
        self.cap_b, self.cap_r = int(bricks), max(1, int(records))
        self.cap_l, self.cap_s = max(1, int(links)), max(1, int(sides))
        self.S = self.cap_b * BC
        dt = torch.float16 if batch.storage == "fp16" else torch.float32
        z = lambda n, d=torch.float32: torch.zeros(n, dtype=d, device=dev)
        self.f = z(self.cap_b * BLOCK, dt)
        self.nbr = torch.full((self.cap_b * 27,), -1, dtype=torch.int32, device=dev)
        self.bidx = torch.full((self.cap_b,), -1, dtype=torch.int32, device=dev)
        self.om_brick = z(self.cap_b, torch.uint8)
        self.om = z(self.S)
        self.mom = BrickMoments(self.S, batch.storage == "fp16", dev)
        self.lmask = z(self.cap_r * BC, torch.int32)
        self.smask = z(self.cap_r * BC, torch.int32)
        self.qlink = z(self.cap_l + 1)                     # one element of padding: the kernels read ahead
        self.qstart = z(self.cap_r, torch.int32)
        self.uin = z(self.cap_r * 3 * BC)
        self.side = z(self.cap_s + 1)
        self.sstart = z(self.cap_r, torch.int32)
        self.loff = z(self.cap_r * BC, torch.int32)
        self.ldir = z(self.cap_r * 2, torch.int32)
        self.n_open = z(batch.max_cases, torch.int32)       # flow-measuring cases of this arena
        self.bricks = _Ranges(self.cap_b)
        self.records = _Ranges(self.cap_r if records > 0 else 0)
        self.links = _Ranges(self.cap_l)
        self.sides = _Ranges(self.cap_s)
        self.cases = []
        self.interior = self.boundary = self.brec = None
        self.n_interior = self.n_boundary = 0
        self.flux = None
        self.dirty = True

    def nbytes(self):
        return arena_bytes(self.cap_b, self.cap_r, self.batch.storage, self.cap_l, self.cap_s)

    def _pools(self, lay):
        return ((self.bricks, int(lay["n_bricks"])), (self.records, int(lay["n_boundary"])),
                (self.links, int(lay["n_links"])), (self.sides, int(lay["n_side"])))

    def place(self, lay):
        """(first brick, first record, first link, first side entry) for the layout, or None."""
        got = []
        for pool, n in self._pools(lay):
            s = pool.alloc(n)
            if s is None:
                for (p2, n2), s2 in zip(self._pools(lay), got):
                    p2.release(s2, n2)
                return None
            got.append(s)
        return tuple(got)

    def release(self, lay, spot):
        for (pool, n), s in zip(self._pools(lay), spot):
            pool.release(s, n)

    def rebuild_lists(self):
        """Device lists from the cases now in the arena. The list buffers have power-of-two capacities and hold skip
        entries past the live ones; the launches cover the live entries (``n_interior``, ``n_boundary``). Returns
        True when a buffer had to grow."""
        torch = _torch()
        dev = self.batch.dev
        shift = _native.CS_CASE_SHIFT
        skip = np.uint32(_native.CS_LIST_SKIP)
        inner = [(c.lay["interior"].astype(np.uint32) + np.uint32(c.b0)) | np.uint32(c.cid << shift)
                 for c in self.cases]
        inner = np.concatenate(inner) if inner else np.zeros(0, np.uint32)
        # boundary list: live entries only, with the record of each entry beside it (brec), so that free records
        # between cases cost no launched workgroups
        bnd, brec = [np.zeros(0, np.uint32)], [np.zeros(0, np.int32)]
        for c in self.cases:
            nbd = int(c.lay["n_boundary"])
            if nbd:
                bnd.append((c.lay["boundary"].astype(np.uint32) + np.uint32(c.b0)) | np.uint32(c.cid << shift))
                brec.append(np.arange(c.r0, c.r0 + nbd, dtype=np.int32))
        bnd, brec = np.concatenate(bnd), np.concatenate(brec)
        grew = False

        def put(buf, data, dtype, fill):
            nonlocal grew
            cap = _pow2(len(data))
            if buf is None or buf.numel() < cap:
                buf = torch.empty(cap, dtype=dtype, device=dev)
                grew = True
            host = np.full(buf.numel(), fill, data.dtype)
            host[:len(data)] = data
            self.batch._upload(buf, host.view(np.int32) if data.dtype == np.uint32 else host)
            return buf

        self.interior = put(self.interior, inner, torch.int32, skip)
        self.boundary = put(self.boundary, bnd, torch.int32, skip)
        self.brec = put(self.brec, brec, torch.int32, 0)
        self.n_interior, self.n_boundary = len(inner), len(bnd)
        # flow-measurement lists (cases with openings under measurement); once present, the flow launch stays in the
        # step (cases without measurement have n_open 0 and their workgroups return at once)
        nopen = np.zeros(self.batch.max_cases, np.int32)
        start = np.zeros(self.batch.max_cases + 1, np.int64)
        idx, w, slot = [np.zeros(0, np.int32)], [np.zeros((0, 3), np.float32)], [np.zeros(0, np.int32)]
        pos = 0
        for cid in range(self.batch.max_cases):
            start[cid] = pos
            c = self.batch.case_by_id.get(cid)
            if c is not None and c.arena is self and c.flux is not None:
                idx.append(c.flux_slots)
                w.append(c.flux["w"])
                slot.append(c.flux["slot"])
                pos += len(c.flux_slots)
                nopen[cid] = c.n_open
        start[self.batch.max_cases] = pos
        if pos or self.flux is not None:
            fx = self.flux or {"start": torch.zeros(self.batch.max_cases + 1, dtype=torch.int64, device=dev),
                               "idx": None, "w": None, "slot": None}
            if self.flux is None:
                grew = True
            fx["idx"] = put(fx["idx"], np.concatenate(idx).astype(np.int32), torch.int32, 0)
            fx["slot"] = put(fx["slot"], np.concatenate(slot).astype(np.int32), torch.int32, 0)
            fx["w"] = put(fx["w"], np.concatenate(w).astype(np.float32).reshape(-1), torch.float32, 0.0)
            # chunk sums of the flow kernels: room for the largest case's chunks (a power of two, at most 256)
            need = max([1] + [_native.flux_chunks(int(start[c + 1] - start[c])) for c in range(self.batch.max_cases)])
            if fx.get("work") is None or fx["chunks"] < need:
                fx["chunks"] = min(256, _pow2(need, lo=1))
                fx["work"] = torch.zeros(self.batch.max_cases * self.batch.max_open * fx["chunks"],
                                         dtype=torch.float64, device=dev)
                grew = True
            self.batch._upload(fx["start"], start)
            self.flux = fx
        self.batch._upload(self.n_open, nopen)
        return grew


def _pow2(n, lo=64):
    c = lo
    while c < n:
        c *= 2
    return c


class _PinnedRing:
    """A ring of pinned host memory for host-to-device copies queued on one stream: a region is reused once the copies
    that last read it have completed (their events); arrays larger than the ring get a fresh pinned copy."""

    def __init__(self, nbytes, stream):
        torch = _torch()
        self.stream = stream
        self.buf = torch.empty(int(nbytes), dtype=torch.uint8, pin_memory=True)
        self.np = self.buf.numpy()
        self.pos, self.busy, self.pending = 0, collections.deque(), []

    def stage(self, x):
        torch = _torch()
        x = np.ascontiguousarray(x)
        n = (x.nbytes + 255) // 256 * 256
        cap = self.buf.numel()
        if n > cap:
            return torch.from_numpy(x).pin_memory()
        if self.pos + n > cap:
            self.staged()                  # the regions of this pass get their event before the ring wraps
            self.pos = 0
        lo, hi = self.pos, self.pos + n
        # the busy regions form a queue in the order their copies were queued on the stream, which is the order they
        # complete in: copies already done leave the front; if a region still in flight overlaps the new one, waiting
        # for the last such region's event completes it and every region before it
        busy = self.busy
        while busy and busy[0][2].query():
            busy.popleft()
        last = -1
        for k, (a, b, _) in enumerate(busy):
            if a < hi and lo < b:
                last = k
        if last >= 0:
            busy[last][2].synchronize()
            for _ in range(last + 1):
                busy.popleft()
        self.np[lo:lo + x.nbytes] = x.reshape(-1).view(np.uint8)
        self.pos = hi
        self.pending.append((lo, hi))
        return self.buf[lo:lo + x.nbytes].view(torch.from_numpy(x[:0]).dtype).reshape(x.shape)

    def staged(self):
        """One event on the ring's stream for the regions staged since the last call (after their copies were
        queued)."""
        if self.pending:
            ev = _torch().cuda.Event()
            ev.record(self.stream)
            self.busy.extend((a, b, ev) for a, b in self.pending)
            self.pending = []


class Batch:
    """A set of cases stepped together (see the module docstring).

    storage "fp32" / "fp16"; model, magic, bouzidi, galilean as in ``lbm.Solver`` (shared by all cases). Arenas
    are created on demand with ``arena_bricks`` / ``arena_records`` capacity while the total stays within
    ``budget_bytes`` (default: the device's free memory less 2 GiB). ``staging_bytes``: the pinned host ring that
    uploads go through; the host waits only when it wraps onto copies still queued behind the steps in flight.
    ``result_bytes``: the pinned host pool that ``Case.read_async`` copies results into (a read that does not fit
    takes a pinned allocation of its own). ``keep_arenas``: an arena left without cases stays allocated (a resident
    server), instead of being freed after a synchronisation."""

    def __init__(self, storage="fp32", model="regularised", magic=3.0 / 16.0, bouzidi=True, galilean=None,
                 device=None, arena_bricks=65536, arena_records=16384, budget_bytes=None, max_cases=256,
                 max_open=8, max_harmonics=16, variant=0, staging_bytes=64 << 20, result_bytes=256 << 20,
                 keep_arenas=False):
        torch = _torch()
        if storage not in STORAGES:
            raise ValueError(f"unknown storage {storage!r}")
        if not 1 <= max_cases <= _native.CS_MAX_CASES:
            raise ValueError(f"max_cases {max_cases}: 1 .. {_native.CS_MAX_CASES}")
        self.dev = _require_gpu(device)
        self.storage = storage
        self.model = MODELS[model]
        self.magic = float(magic)
        self.bouzidi = int(bool(bouzidi))
        self.gal = int(True if galilean is None else bool(galilean))
        self.arena_bricks, self.arena_records = int(arena_bricks), int(arena_records)
        if budget_bytes is None:
            free, _ = torch.cuda.mem_get_info(self.dev)
            budget_bytes = max(0, free - 2 * 2 ** 30)
        self.budget = int(budget_bytes)
        self.max_cases, self.max_open, self.max_harm = int(max_cases), int(max_open), int(max_harmonics)
        self.variant = int(variant)
        self.keep_arenas = bool(keep_arenas)
        self._tstream = torch.cuda.Stream(device=self.dev)          # default priority
        self.stream_handle = ctypes.c_void_p(self._tstream.cuda_stream)
        M, K = self.max_cases, self.max_open
        with torch.cuda.stream(self._tstream):
            z = lambda shape, d=torch.float32: torch.zeros(shape, dtype=d, device=self.dev)
            self.tick = z(M, torch.int32)
            self.coef = z((M, 2 * self.max_harm + 1))
            self.nh = z(M, torch.int32)
            self.period = z(M, torch.int32)
            self.ramp = z(M, torch.int32)
            self.drive = z((256, M))           # drives of the steps of a chunk, per case
            self._k = 0
            self.omega = z(M)
            self.force = z((M, 3))
            self.p_out = z((M, K))
            self.q_last = z((M, K))
            self.R = z((M, K))
            self.pref = z(M)
            self.alpha = z(M)
            self._work = z(_native.CS_REDUCE_WORK, torch.float64)
            self._red = z(2, torch.float64)
        self._tstream.synchronize()
        self._ring = _PinnedRing(int(staging_bytes), self._tstream)   # upload staging (batch stream)
        self._cstream, self._bring, self._blob_bytes = None, None, int(staging_bytes)   # packed admissions
        # pinned pool of result reads, in 256-byte units, allocated at the first read (allocating pinned memory per
        # read can wait for the device)
        self._rpool, self._rpool_bytes = None, int(result_bytes)
        self._rranges = _Ranges(self._rpool_bytes // 256)
        self.arenas = []
        self.topo = 0                      # bumped when the launch structure changes (graphs captured again)
        self.case_by_id = {}
        self.cur = 0
        self.steps = 0
        self.graphs = {}
        self._retired = []                 # (graphs, event after their last replay) awaiting destruction
        self.captures, self.capture_s = 0, 0.0
        self._args_key, self._args = None, {}
        self._blocks = None

    def _stage(self, x):
        """A pinned host tensor holding x, in the batch's staging ring (copies on the batch stream)."""
        return self._ring.stage(x)

    def _staged(self):
        """Record one event for the ring regions staged since the last call (after their copies were queued)."""
        self._ring.staged()

    def _blob_to_device(self, buf):
        """A new device tensor holding the uint8 buffer buf (a case's packed arrays), copied on the batch's copy
        stream: the copy does not queue behind the chunks in flight (it writes fresh memory), so its staging region
        frees as soon as it is done. The batch stream waits for it before any launch that follows."""
        torch = _torch()
        if self._cstream is None:
            self._cstream = torch.cuda.Stream(device=self.dev)          # default priority
            self._bring = _PinnedRing(self._blob_bytes, self._cstream)
        src = self._bring.stage(buf)
        with torch.cuda.stream(self._cstream):
            out = src.to(self.dev, non_blocking=True)
        self._bring.staged()
        out.record_stream(self._tstream)
        self._tstream.wait_stream(self._cstream)
        return out

    def _result_buffer(self, shapes):
        """Pinned host float32 tensors of the given shapes for a read's copies, in one region of the result pool, and
        a function that returns the region; (fresh pinned tensors, None) when the pool has no room."""
        torch = _torch()
        if self._rpool is None and self._rpool_bytes:
            self._rpool = torch.empty(self._rpool_bytes, dtype=torch.uint8, pin_memory=True)
        sizes = [(int(np.prod(sh)) * 4 + 255) // 256 for sh in shapes]
        start = self._rranges.alloc(sum(sizes)) if self._rpool is not None else None
        if start is None:
            return [torch.empty(sh, dtype=torch.float32, pin_memory=True) for sh in shapes], None
        out, pos = [], start * 256
        for sh, units in zip(shapes, sizes):
            n = int(np.prod(sh))
            out.append(self._rpool[pos:pos + 4 * n].view(torch.float32).view(sh))
            pos += units * 256
        return out, lambda: self._rranges.release(start, sum(sizes))

    def _upload(self, dst, x, staged=True):
        """Asynchronous host-to-device copy into dst on the batch stream: ordered after the work already queued and
        before the next launches, without blocking the host. staged=False leaves the ring event to a later
        ``_staged()`` (one event for a group of uploads)."""
        torch = _torch()
        x = np.asarray(x).astype(torch.empty(0, dtype=dst.dtype).numpy().dtype, copy=False).reshape(dst.shape)
        src = self._stage(x)
        with torch.cuda.stream(self._tstream):
            dst.copy_(src, non_blocking=True)
        if staged:
            self._staged()

    def _to_device(self, x, staged=True):
        """A new device tensor holding x, copied asynchronously on the batch stream."""
        torch = _torch()
        src = self._stage(x)
        with torch.cuda.stream(self._tstream):
            out = src.to(self.dev, non_blocking=True)
        if staged:
            self._staged()
        return out

    def _scatter_rows(self, dst, rows, data):
        """Rows `rows` (host int array) of the 2-D device view dst set to data (len(rows), dst.shape[1]): one upload
        of the rows and their indices, then an index copy on the batch stream."""
        torch = _torch()
        if not len(rows):
            return
        idx = self._to_device(np.asarray(rows, np.int64), staged=False)
        src = self._to_device(np.ascontiguousarray(data, np.float32), staged=False)
        with torch.cuda.stream(self._tstream):
            dst.index_copy_(0, idx, src)
        self._staged()

    # --- capacity -------------------------------------------------------------------------------------------
    def nbytes(self):
        return sum(a.nbytes() for a in self.arenas)

    def _arena_shape(self, lay):
        nb = min(max(self.arena_bricks, int(lay["n_bricks"])), MAX_BRICKS)
        nr = max(self.arena_records, int(lay["n_boundary"]))
        nl = max(LINKS_PER_RECORD * self.arena_records, int(lay["n_links"]))
        ns = max(SIDES_PER_RECORD * self.arena_records, int(lay["n_side"]))
        return nb, nr, nl, ns

    def arena_bytes_for(self, lay):
        """Bytes of the arena this batch would open for a case of this layout."""
        nb, nr, nl, ns = self._arena_shape(lay)
        return arena_bytes(nb, nr, self.storage, nl, ns)

    def _new_arena(self, lay):
        nb, nr, nl, ns = self._arena_shape(lay)
        if lay["n_bricks"] > nb:
            raise ValueError(f"case of {lay['n_bricks']} bricks exceeds one arena ({MAX_BRICKS} bricks)")
        if self.nbytes() + arena_bytes(nb, nr, self.storage, nl, ns) > self.budget:
            return None
        with _torch().cuda.stream(self._tstream):
            a = Arena(self, nb, nr, nl, ns)
        self.arenas.append(a)
        self.topo += 1
        return a

    def reserve(self):
        """Open one arena of the default capacity now (a serving loop's start-up), so that the first admission does
        not pay for allocating and clearing it. Nothing happens when an arena exists or the budget has no room."""
        if self.arenas:
            return
        nb, nr = min(self.arena_bricks, MAX_BRICKS), self.arena_records
        nl, ns = LINKS_PER_RECORD * nr, SIDES_PER_RECORD * nr
        if arena_bytes(nb, nr, self.storage, nl, ns) > self.budget:
            return
        with _torch().cuda.stream(self._tstream):
            self.arenas.append(Arena(self, nb, nr, nl, ns))
        self.topo += 1
        self.synchronize()

    def can_admit(self, lay):
        """Whether a case of this layout fits now (free case slot and room in an arena or the budget)."""
        if len(self.case_by_id) >= self.max_cases:
            return False
        for a in self.arenas:
            spot = a.place(lay)
            if spot is not None:
                a.release(lay, spot)
                return True
        nb, nr, nl, ns = self._arena_shape(lay)
        # arenas kept without cases (keep_arenas) give their room back when a new one is needed
        held = sum(a.nbytes() for a in self.arenas if a.cases)
        return held + arena_bytes(nb, nr, self.storage, nl, ns) <= self.budget

    def _drop_empty_arenas(self):
        """Free the arenas without cases (after a synchronisation); True when there was one."""
        empty = [a for a in self.arenas if not a.cases]
        if not empty:
            return False
        self.synchronize()
        for a in empty:
            self.arenas.remove(a)
        self.topo += 1
        return True

    # --- cases ----------------------------------------------------------------------------------------------
    def add(self, lat, nu_lu, layout=None, packed=None):
        """Admit a lattice as a new case at rest (f~ = 0, u = 0); returns its ``Case`` handle. Raises BatchFull.
        packed: the case's upload arrays from ``pack_admission`` (prepared off this thread), so that admitting costs
        one staging copy and one host-to-device copy here."""
        from . import bricks
        lay = layout if layout is not None else bricks.build(lat)
        if len(self.case_by_id) >= self.max_cases:
            raise BatchFull("no free case slot")
        cid = min(set(range(self.max_cases)) - set(self.case_by_id))
        spot = None
        for a in self.arenas:
            p = a.place(lay)
            if p is not None:
                spot = (a, p)
                break
        if spot is None:
            a = self._new_arena(lay)
            if a is None and self._drop_empty_arenas():
                a = self._new_arena(lay)
            if a is None:
                raise BatchFull("HBM budget exhausted")
            p = a.place(lay)
            if p is None:
                raise BatchFull("case does not fit a fresh arena")
            spot = (a, p)
        c = Case(self, cid, lat, nu_lu, lay, spot[0], spot[1], packed=packed)
        self.case_by_id[cid] = c
        spot[0].cases.append(c)
        self._topology_changed(spot[0])
        return c

    def remove(self, case):
        """Release a case's bricks, records and slot. Stream-ordered, without waiting for the device: the launches and
        copies already queued (the case's last steps, reads of its results) run before anything that reuses its
        ranges or its slot (the uploads of a new case, the rebuilt lists). Only an arena left empty is freed after a
        synchronisation."""
        torch = _torch()
        a = case.arena
        a.cases.remove(case)
        a.release(case.lay, case.spot)
        with torch.cuda.stream(self._tstream):
            self.period[case.cid:case.cid + 1].zero_()     # a fill kernel: no wait for the queued steps
        del self.case_by_id[case.cid]
        case.arena = None
        self._topology_changed(a)
        if not a.cases and not self.keep_arenas:
            self.synchronize()
            self.arenas.remove(a)
            self.topo += 1

    def _topology_changed(self, arena):
        """The arena's membership changed: its lists are rebuilt before the next launch."""
        arena.dirty = True

    def _flush(self):
        """Rebuild the lists of the arenas whose membership changed (stream-ordered after the work in flight)."""
        with _torch().cuda.stream(self._tstream):
            for ar in self.arenas:
                if ar.dirty:
                    if ar.rebuild_lists():
                        self.topo += 1
                    ar.dirty = False


    # --- stepping -------------------------------------------------------------------------------------------
    def _drive_ptr(self):
        """Device address of the per-case drives of the step being launched (table row self._k)."""
        return self.drive.data_ptr() + 4 * self.max_cases * self._k

    def _clock_table(self, n):
        """Every case's clock for the next n steps at once; rows 0..n-1 of the drive table."""
        A = _native.ClockTableArgs()
        A.tick, A.coef, A.nh = self.tick.data_ptr(), self.coef.data_ptr(), self.nh.data_ptr()
        A.period, A.ramp, A.drive = self.period.data_ptr(), self.ramp.data_ptr(), self.drive.data_ptr()
        A.n_cases, A.coef_stride, A.n_steps, A.drive_stride = self.max_cases, self.coef.shape[1], int(n), \
            self.max_cases
        _native.call("cs_clock_table", ctypes.byref(A), self.stream_handle)

    def _clock(self):
        """The clocks of one step (eager launches)."""
        self._clock_table(1)
        self._k = 0

    def _grow_table(self, n):
        if n > self.drive.shape[0]:
            torch = _torch()
            self.synchronize()
            self.graphs = {}
            with torch.cuda.stream(self._tstream):
                self.drive = torch.zeros((n, self.max_cases), dtype=torch.float32, device=self.dev)
            self.synchronize()

    def step_args(self, arena, a, boundary):
        A = _native.BricksStepArgs()
        lst = arena.boundary if boundary else arena.interior
        A.f, A.nbr, A.list = arena.f.data_ptr(), arena.nbr.data_ptr(), lst.data_ptr()
        A.n_list = arena.n_boundary if boundary else arena.n_interior    # the live extent of the list
        A.bidx, A.lmask, A.smask = arena.bidx.data_ptr(), arena.lmask.data_ptr(), arena.smask.data_ptr()
        A.qlink, A.qstart, A.uin = arena.qlink.data_ptr(), arena.qstart.data_ptr(), arena.uin.data_ptr()
        A.side, A.sstart = arena.side.data_ptr(), arena.sstart.data_ptr()
        A.loff, A.ldir = arena.loff.data_ptr(), arena.ldir.data_ptr()
        A.om_cell, A.om_brick, A.omega = arena.om.data_ptr(), arena.om_brick.data_ptr(), 1.0
        A.p_out, A.drive = self.p_out.data_ptr(), self._drive_ptr()
        A.vel_old, A.vel, A.rho = arena.mom.step_ptrs(a)
        A.n_slots, A.storage = arena.S, STORAGES[self.storage]
        A.magic, A.bouzidi, A.model, A.galilean = self.magic, self.bouzidi, self.model, self.gal
        A.parity, A.boundary = a, int(boundary)
        A.case_omega, A.case_force = self.omega.data_ptr(), self.force.data_ptr()
        A.p_out_stride, A.variant = self.max_open, self.variant
        A.brec = arena.brec.data_ptr() if boundary else None
        return A

    def _flux(self, arena, a):
        fx = arena.flux
        A = _native.FluxCasesArgs()
        A.idx, A.w, A.slot, A.start = fx["idx"].data_ptr(), fx["w"].data_ptr(), fx["slot"].data_ptr(), \
            fx["start"].data_ptr()
        A.vel = arena.mom.velocity(1 - a).data_ptr()
        A.vel_elem, A.vel_comp = arena.mom.velocity_addressing()
        A.R, A.drive, A.p_out, A.q_last = self.R.data_ptr(), self._drive_ptr(), self.p_out.data_ptr(), \
            self.q_last.data_ptr()
        A.n_open, A.pref, A.alpha = arena.n_open.data_ptr(), self.pref.data_ptr(), self.alpha.data_ptr()
        A.n_cases, A.open_stride = self.max_cases, self.max_open
        A.work, A.work_chunks = fx["work"].data_ptr(), fx["chunks"]
        _native.call("cs_flux_cases", ctypes.byref(A), self.stream_handle)

    def launch_parts(self):
        """(arena index, boundary flag, cells) of the step launches of one step, in order."""
        out = []
        for k, ar in enumerate(self.arenas):
            ni = sum(int(c.lay["n_interior"]) * BC for c in ar.cases)
            nf = sum(c.N for c in ar.cases)
            out += [(k, False, ni), (k, True, nf - ni)]
        return out

    def step_part(self, k, boundary, a):
        self._flush()
        ar = self.arenas[k]
        A = self.step_args(ar, a, boundary)
        if A.n_list:
            _native.call("cs_bricks_step", ctypes.byref(A), self.stream_handle)

    def _step_pairs(self, a):
        """[(arena, boundary args, interior args)] of the step of parity a, with the drive pointer of table row
        self._k: one paired launch per arena with bricks (cs_bricks_step_pair; an empty list is skipped there). The
        argument blocks are kept while the arenas, their list extents and the variant stay the same (graph captures
        then cost little more than the launches themselves)."""
        key = (self.topo, self.variant, tuple((id(ar), ar.n_interior, ar.n_boundary,
                                               ar.interior.data_ptr() if ar.interior is not None else 0,
                                               ar.boundary.data_ptr() if ar.boundary is not None else 0,
                                               ar.brec.data_ptr() if ar.brec is not None else 0)
                                              for ar in self.arenas))
        if key != self._args_key:
            self._args_key, self._args = key, {}
        got = self._args.get(a)
        if got is None:
            got = [(ar, self.step_args(ar, a, True), self.step_args(ar, a, False)) for ar in self.arenas]
            got = self._args[a] = [x for x in got if x[1].n_list or x[2].n_list]
        drive = self._drive_ptr()
        for _, Ab, Ai in got:
            Ab.drive = Ai.drive = drive
        return got

    def _native_blocks(self):
        """(argument blocks of both parities, pairs per parity) for cs_bricks_run_pairs, with the drive pointers at
        table row 0; kept while the step argument blocks and the drive table stay the same."""
        self._k = 0
        b0, b1 = self._step_pairs(0), self._step_pairs(1)
        key = (self._args_key, self.drive.data_ptr())
        if self._blocks is None or self._blocks[0] != key:
            assert [id(ar) for ar, _, _ in b0] == [id(ar) for ar, _, _ in b1]
            n = len(b0)
            blocks = [A for _, Ab, Ai in b0 + b1 for A in (Ab, Ai)]
            arr = (_native.BricksStepArgs * max(1, 4 * n))(*blocks)
            self._blocks = (key, arr, n)
        return self._blocks[1:]

    def _steps(self, n, a):
        """The launches of n steps from parity a, drives at table rows 0..n-1: one native call (cs_bricks_run_pairs:
        one paired launch per arena and step); the per-step Python launches when cases measure flows (their launch
        follows each step)."""
        if any(ar.flux is not None for ar in self.arenas):
            for k in range(n):
                self._k = k
                self._launch(a)
                a = 1 - a
            return
        arr, npairs = self._native_blocks()
        if npairs:
            _native.call("cs_bricks_run_pairs", arr, npairs, int(n), int(a), self.max_cases, self.stream_handle)
        self._k = n - 1

    def _launch(self, a):
        """The launches of one step of parity a (drives at table row self._k): one paired launch per arena (the
        brick lists of a step are independent, as in lbm._BricksBackend.step; arenas share nothing), then the flow
        launches, which read the step's velocities."""
        for ar, Ab, Ai in self._step_pairs(a):
            _native.call("cs_bricks_step_pair", ctypes.byref(Ab), ctypes.byref(Ai), self.stream_handle)
        for ar in self.arenas:
            if ar.flux is not None:
                self._flux(ar, a)

    def run(self, steps, chunk=200):
        """Advance every case `steps` steps: graphs of `chunk` steps (made even), remainder step by step."""
        if not self.case_by_id:
            return
        self._flush()
        chunk += chunk % 2
        full, rest = divmod(int(steps), chunk)
        if full:
            self._grow_table(chunk)
            # the launches cover the live extent of each list (skipped entries still cost a workgroup dispatch each),
            # so a graph is captured for each set of extents; graphs of other extents are retired and destroyed once
            # the device has passed their last replay
            sizes = tuple((ar.n_interior, ar.n_boundary) for ar in self.arenas)
            key = (self.cur, chunk, self.variant, self.topo, sizes)
            self._purge_retired()
            if key not in self.graphs:
                stale = [k for k in self.graphs if k[3:] != key[3:]]
                if stale:
                    ev = _torch().cuda.Event()
                    ev.record(self._tstream)
                    self._retired.append(([self.graphs.pop(k) for k in stale], ev))
                tc = time.perf_counter()
                with _native.capture(self.stream_handle) as cap:
                    self._clock_table(chunk)
                    self._steps(chunk, self.cur)
                self.graphs[key] = cap.graph
                self.capture_s += time.perf_counter() - tc
                self.captures += 1
            g = self.graphs[key]
            for _ in range(full):
                g.launch(self.stream_handle)
            self._k = chunk - 1
        for _ in range(rest):
            self._clock()
            self._launch(self.cur)
            self.cur = 1 - self.cur
        self.steps += int(steps)
        for c in self.case_by_id.values():
            c.steps += int(steps)

    def _purge_retired(self):
        self._retired = [(gs, ev) for gs, ev in self._retired if not ev.query()]

    def synchronize(self):
        self._tstream.synchronize()

    # --- autotuning -----------------------------------------------------------------------------------------
    def variants(self):
        return launch_variants()

    def autotune(self, steps=20, repeats=3):
        """Time every launch variant on this batch as run() steps it (HIP events on the batch stream around replays
        of a captured graph of `steps` steps, made even: one untimed replay, then the best of `repeats`), keep the
        fastest, and restore the state the timing runs changed. Returns {variant: ms/step}."""
        torch = _torch()
        self._flush()
        snap = self._snapshot()
        counters = (self.cur, self.steps, {cid: c.steps for cid, c in self.case_by_id.items()})
        steps = max(2, int(steps) + int(steps) % 2)
        times = {}
        try:
            for var in self.variants():
                self.variant = var["variant"]
                self.run(steps, chunk=steps)
                best = None
                for _ in range(repeats):
                    e0 = torch.cuda.Event(enable_timing=True)
                    e1 = torch.cuda.Event(enable_timing=True)
                    e0.record(self._tstream)
                    self.run(steps, chunk=steps)
                    e1.record(self._tstream)
                    self.synchronize()
                    ms = e0.elapsed_time(e1) / steps
                    best = ms if best is None else min(best, ms)
                times[var["variant"]] = best
        finally:
            self._restore(snap)
            self.cur, self.steps = counters[0], counters[1]
            for cid, n in counters[2].items():
                self.case_by_id[cid].steps = n
        self.variant = min(times, key=times.get)
        self.graphs = {}
        return times

    def _snapshot(self):
        self.synchronize()
        with _torch().cuda.stream(self._tstream):
            s = {"tables": [t.clone() for t in (self.tick, self.drive, self.p_out, self.q_last)],
                 "arenas": [[ar.f.clone(), ar.side.clone()] + [t.clone() for t in ar.mom.tensors()]
                            for ar in self.arenas]}
        self.synchronize()
        return s

    def _restore(self, s):
        self.synchronize()
        with _torch().cuda.stream(self._tstream):
            for dst, src in zip((self.tick, self.drive, self.p_out, self.q_last), s["tables"]):
                dst.copy_(src)
            for ar, saved in zip(self.arenas, s["arenas"]):
                for dst, src in zip([ar.f, ar.side] + ar.mom.tensors(), saved):
                    dst.copy_(src)
        self.synchronize()


def sponge_slots(lat, lay, tau0, length, tau_max, kinds=("inlet", "outlet")):
    """Host part of set_sponge: omega per slot of the layout's bricks and the flags of the bricks whose cells
    differ from 1 / tau0."""
    tau = sponge_tau(lat, tau0, length, tau_max, kinds)
    omega = np.float32(1.0 / tau0)
    nb = int(lay["n_bricks"])
    slots = np.full(nb * BC, omega, np.float32)
    slots[lay["slot_of_cell"]] = (1.0 / tau).astype(np.float32)
    own = (slots.reshape(nb, BC) != omega).any(1).astype(np.uint8)
    return slots, own


class PackedArrays:
    """Host arrays packed into one buffer (each at a 256-byte aligned offset) and the device views of its copy."""

    def __init__(self, arrays):
        self.offs, total = {}, 0
        arrays = {k: np.ascontiguousarray(a) for k, a in arrays.items()}
        for k, a in arrays.items():
            self.offs[k] = (total, a.dtype, a.shape)
            total += (a.nbytes + 255) // 256 * 256
        self.buf = np.empty(max(total, 256), np.uint8)
        for k, a in arrays.items():
            o = self.offs[k][0]
            self.buf[o:o + a.nbytes] = a.reshape(-1).view(np.uint8)

    def __contains__(self, key):
        return key in self.offs

    def view(self, dev, key):
        """The array `key` in the device copy `dev` of the buffer (a view: dtype and shape of the host array)."""
        torch = _torch()
        o, dt, shape = self.offs[key]
        n = int(np.prod(shape)) * dt.itemsize
        tdt = torch.from_numpy(np.empty(0, dt)).dtype
        return dev[o:o + n].view(tdt).view(shape)


def pack_admission(lay, rec=None):
    """The arrays a case's admission uploads, in their device types and packed into one buffer: the layout's arrays
    (Case) and the per-slot omegas of a recorded sponge (``SetupRecorder``). Runs on a serving worker thread, so that
    the serving thread's admission is one staging copy, one host-to-device copy and device-side copies into place."""
    nb, nr = int(lay["n_bricks"]), int(lay["n_boundary"])
    a = {"slot": np.asarray(lay["slot_of_cell"], np.int32), "nbr": np.asarray(lay["nbr"], np.int32).reshape(-1),
         "bidx": np.asarray(lay["bidx"], np.int32)}
    if nr:
        a["lmask"] = lay["lmask"].view(np.int32).reshape(-1)
        a["smask"] = lay["smask"].view(np.int32).reshape(-1)
        inl = np.flatnonzero((lay["lmask"] & np.uint32(1 << 27)).any(axis=1))
        if len(inl):
            a["inl_idx"] = inl.astype(np.int64)
            a["inl_rows"] = np.asarray(lay["uin"], np.float32).reshape(nr, 3 * BC)[inl]
        a["qstart"] = np.asarray(lay["qstart"][:-1], np.int32)
        a["sstart"] = np.asarray(lay["sstart"][:-1], np.int32)
        a["loff"] = lay["loff"].view(np.int32).reshape(-1)
        a["ldir"] = lay["ldir"].view(np.int32).reshape(-1)
        if int(lay["n_links"]):
            a["qlink"] = np.asarray(lay["qlink"][:int(lay["n_links"])], np.float32)
    for name, args, _ in (rec.calls if rec is not None else []):
        if name == "set_omega_slots":
            slots, own = args
            rows = np.flatnonzero(own)
            a["om_own"] = np.asarray(own, np.uint8)
            if len(rows):
                a["om_rows"] = rows.astype(np.int64)
                a["om_data"] = np.asarray(slots, np.float32).reshape(nb, BC)[rows]
    return PackedArrays(a)


class SetupRecorder:
    """Stands in for a Case while a request is prepared off the serving thread: records the setter calls, and
    computes the sponge's per-slot omega there. ``replay(case)`` applies them to the admitted case."""

    def __init__(self, lat, nu_lu, lay):
        self.lat, self.lay, self.tau = lat, lay, 3.0 * nu_lu + 0.5
        self.calls = []
        self.packed = None                 # pack_admission(lay, self), when prepared

    def set_drive(self, *a, **k):
        self.calls.append(("set_drive", a, k))

    def set_force(self, *a, **k):
        self.calls.append(("set_force", a, k))

    def set_outlet_pressure(self, *a, **k):
        self.calls.append(("set_outlet_pressure", a, k))

    def set_outlets(self, *a, **k):
        self.calls.append(("set_outlets", a, k))

    def set_sponge(self, length, tau_max, kinds=("inlet", "outlet")):
        self.calls.append(("set_omega_slots", sponge_slots(self.lat, self.lay, self.tau, length, tau_max, kinds), {}))

    def replay(self, case):
        for name, a, k in self.calls:
            if name == "set_omega_slots" and case.packed_sponge:
                continue                   # applied at admission from the packed arrays
            getattr(case, name)(*a, **k)


class PendingFields:
    """(pressure, velocity) of a case on their way to pinned host memory (``Case.read_async``)."""

    def __init__(self, event, rho, u, device_copies, release=None):
        self.event, self._rho, self._u = event, rho, u
        self._keep = device_copies           # the gathered device tensors stay referenced until the copy is done
        self._release = release              # returns the pinned region to the batch's result pool

    def ready(self):
        return self.event.query()

    def result(self):
        """Wait for the copies; (pressure, velocity) as ``Case.fields`` returns them, and whether every density
        is finite (``Case.ok`` over the fluid cells). The pinned region goes back to the pool (results copied out)."""
        self.event.synchronize()
        self._keep = None
        rho, u = self._rho.numpy(), self._u.numpy()
        if self._release is not None:
            u = u.copy()
            out = (rho / 3.0, u), bool(np.isfinite(rho).all())
            self._rho = self._u = None
            self._release()
            self._release = None
            return out
        return (rho / 3.0, u), bool(np.isfinite(rho).all())


class Case:
    """One lattice inside a Batch, with the lbm.Solver surface (lbm.py:346-455) and the resistance-outlet flow
    measurement of simulate.py (ArterySolver)."""

    def __init__(self, batch, cid, lat, nu_lu, lay, arena, spot, packed=None):
        torch = _torch()
        b0, r0, q0, s0 = spot
        self.batch, self.cid, self.lat, self.lay, self.arena = batch, cid, lat, lay, arena
        self.spot, self.b0, self.r0 = spot, b0, r0
        self.N = int(lay["n_cells"])
        self.tau = 3.0 * nu_lu + 0.5
        self.steps = 0
        self.flux = None
        self.n_open = 0
        nb, nr = int(lay["n_bricks"]), int(lay["n_boundary"])
        # Every upload below is asynchronous on the batch stream: queued after the chunks in flight (which may still
        # step a case that left this range) and before the next launches. The layout's arrays go up as they are and
        # the arena offsets are added on the device (less host work per admission); one ring event covers them all.
        # With packed arrays (pack_admission, prepared on a worker thread) they go up as one buffer and are copied
        # into place on the device; the device buffer stays with the case (its slot table is a view of it).
        self.packed_sponge = packed is not None and "om_own" in packed
        dev = batch._blob_to_device(packed.buf) if packed is not None else None
        self._packed_dev = dev

        def up(dst, x, key=None):
            if dev is not None and key is not None:
                with torch.cuda.stream(batch._tstream):
                    dst.copy_(packed.view(dev, key))
            else:
                batch._upload(dst, x, staged=False)

        with torch.cuda.stream(batch._tstream):
            # slot of each cell (int32: the gathers index with it too)
            self.slot32 = (packed.view(dev, "slot") if dev is not None
                           else batch._to_device(lay["slot_of_cell"], staged=False))
            self.slot32.add_(b0 * BC)
            self.slot_dev = self.slot32
            nbr = arena.nbr[b0 * 27:(b0 + nb) * 27]
            up(nbr, lay["nbr"].reshape(-1), "nbr")
            nbr.add_(nbr.ge(0).int().mul_(b0))                      # neighbour bricks; -1 (absent) stays
            bidx = arena.bidx[b0:b0 + nb]
            up(bidx, lay["bidx"], "bidx")
            bidx.add_(bidx.ge(0).int().mul_(r0))                    # boundary records; -1 (interior) stays
            arena.om_brick[b0:b0 + nb].zero_()
            arena.om[b0 * BC:(b0 + nb) * BC].zero_()
            arena.f[b0 * BLOCK:(b0 + nb) * BLOCK].zero_()
            arena.mom.zero(b0 * BC, (b0 + nb) * BC)
            if nr:
                up(arena.lmask[r0 * BC:(r0 + nr) * BC], lay["lmask"].view(np.int32).reshape(-1), "lmask")
                up(arena.smask[r0 * BC:(r0 + nr) * BC], lay["smask"].view(np.int32).reshape(-1), "smask")
                # inlet profiles: only the records with inlet cells read theirs
                if dev is not None:
                    if "inl_idx" in packed:
                        idx = packed.view(dev, "inl_idx").add_(r0)
                        arena.uin.view(-1, 3 * BC).index_copy_(0, idx, packed.view(dev, "inl_rows"))
                else:
                    inl = np.flatnonzero((lay["lmask"] & np.uint32(1 << 27)).any(axis=1))
                    batch._scatter_rows(arena.uin.view(-1, 3 * BC), inl + r0, lay["uin"].reshape(nr, 3 * BC)[inl])
                qs, ss = arena.qstart[r0:r0 + nr], arena.sstart[r0:r0 + nr]
                up(qs, lay["qstart"][:-1], "qstart")
                qs.add_(q0)
                up(ss, lay["sstart"][:-1], "sstart")
                ss.add_(s0)
                up(arena.loff[r0 * BC:(r0 + nr) * BC], lay["loff"].view(np.int32).reshape(-1), "loff")
                up(arena.ldir[r0 * 2:(r0 + nr) * 2], lay["ldir"].view(np.int32).reshape(-1), "ldir")
                if lay["n_links"]:
                    up(arena.qlink[q0:q0 + int(lay["n_links"])], lay["qlink"][:int(lay["n_links"])], "qlink")
                if lay["n_side"]:
                    arena.side[s0:s0 + int(lay["n_side"])].zero_()
            # per-case table entries by fill kernels (an indexed assignment of a Python number copies it from host
            # memory, which waits for the work queued on the stream)
            c = cid
            batch.tick[c:c + 1].zero_()
            batch.drive[:, c].zero_()
            batch.omega[c:c + 1].fill_(float(np.float32(1.0 / self.tau)))
            batch.force[c].zero_()
            batch.p_out[c].zero_()
            batch.q_last[c].zero_()
            if self.packed_sponge:
                # the recorded sponge: per-slot omegas of the flagged bricks and the flags
                if "om_rows" in packed:
                    rows = packed.view(dev, "om_rows").add_(b0)
                    arena.om.view(-1, BC).index_copy_(0, rows, packed.view(dev, "om_data"))
                arena.om_brick[b0:b0 + nb].copy_(packed.view(dev, "om_own"))
        batch._staged()
        self.set_drive([1.0])

    # --- reference surface ----------------------------------------------------------------------------------
    @property
    def cur(self):
        return self.batch.cur

    def _set(self, tensor, value):
        self.batch._upload(tensor, np.asarray(value))

    def set_drive(self, coef, period=1, ramp=0):
        """As lbm.Solver.set_drive (lbm.py:384-392): the case's clock restarts from its current tick."""
        coef = np.asarray(coef, np.float32)
        if len(coef) % 2 == 0:
            coef = np.append(coef, 0.0).astype(np.float32)
        nh = (len(coef) - 1) // 2
        if nh > self.batch.max_harm:
            raise ValueError(f"{nh} harmonics > max_harmonics {self.batch.max_harm}")
        row = np.zeros(self.batch.coef.shape[1], np.float32)
        row[:len(coef)] = coef
        b, c = self.batch, self.cid
        self._set(b.coef[c], row)
        self._set(b.nh[c:c + 1], [nh])
        self._set(b.ramp[c:c + 1], [int(ramp)])
        self._set(b.period[c:c + 1], [int(period)])

    def set_force(self, F):
        self._set(self.batch.force[self.cid], [float(np.float32(x)) for x in F])

    def set_sponge(self, length, tau_max, kinds=("inlet", "outlet")):
        """As lbm.Solver.set_sponge (lbm.py:398-416)."""
        self.set_omega_slots(*sponge_slots(self.lat, self.lay, self.tau, length, tau_max, kinds))

    def set_omega_slots(self, slots, own):
        """Per-slot omega of the case's bricks and the bricks that need it (from sponge_slots). Only the flagged
        bricks' slots are uploaded: the step reads a brick's per-slot omegas only when its flag is set."""
        nb, ar, b0 = int(self.lay["n_bricks"]), self.arena, self.b0
        rows = np.flatnonzero(own)
        self.batch._scatter_rows(ar.om.view(-1, BC), rows + b0, np.asarray(slots, np.float32).reshape(nb, BC)[rows])
        self.batch._upload(ar.om_brick[b0:b0 + nb], own)

    def set_outlet_pressure(self, p_lu):
        p = np.zeros(self.batch.max_open, np.float32)
        v = np.asarray(p_lu, np.float32).reshape(-1)
        if len(v) > self.batch.max_open:
            raise ValueError(f"{len(v)} openings > max_open {self.batch.max_open}")
        p[:len(v)] = v
        self._set(self.batch.p_out[self.cid], p)

    def set_outlets(self, flux, R_lu, pref_lu, alpha):
        """Resistance outlets with per-step flow measurement (simulate.py:93-114, as artery.ArterySolver)."""
        n = len(R_lu)
        if not 1 <= n <= self.batch.max_open:
            raise ValueError(f"{n} openings: 1 .. {self.batch.max_open}")
        R = np.zeros(self.batch.max_open, np.float32)
        R[:n] = R_lu
        self._set(self.batch.R[self.cid], R)
        self._set(self.batch.pref[self.cid:self.cid + 1], [pref_lu])
        self._set(self.batch.alpha[self.cid:self.cid + 1], [alpha])
        self.flux = {"w": np.ascontiguousarray(flux["w"], np.float32),
                     "slot": np.ascontiguousarray(flux["slot"], np.int32)}
        self.flux_slots = (self.lay["slot_of_cell"][np.asarray(flux["idx"])].astype(np.int64) +
                           self.b0 * BC).astype(np.int32)
        self.n_open = n
        self.batch._topology_changed(self.arena)

    def flows(self):
        self.batch.synchronize()
        return self.batch.q_last[self.cid, :max(1, self.n_open)].cpu().numpy()

    def opening_pressures(self):
        self.batch.synchronize()
        n = max(1, self.n_open, self.lat["n_outlets"])
        return self.batch.p_out[self.cid, :n].cpu().numpy()

    def read_async(self):
        """Queue the copy of the case's (pressure, velocity) to pinned host memory behind the work in flight and
        return a ``PendingFields`` (``ready()``, ``result()``) without waiting. The case may leave the batch right
        after: the copies run before anything that reuses its ranges."""
        torch = _torch()
        bt = self.batch
        with torch.cuda.stream(bt._tstream):
            rho, u = self.arena.mom.gather(bt.cur, self.slot_dev)
            (rh, uh), release = bt._result_buffer([tuple(rho.shape), tuple(u.shape)])
            rh.copy_(rho, non_blocking=True)
            uh.copy_(u, non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(bt._tstream)
        return PendingFields(ev, rh, uh, (rho, u), release)

    def fields(self):
        """(pressure, velocity) per cell in the lattice's order."""
        self.batch.synchronize()
        with _torch().cuda.stream(self.batch._tstream):
            ar, cur = self.arena, self.batch.cur
            rho, u = ar.mom.gather(cur, self.slot_dev)
            out = rho.cpu().numpy() / 3.0, u.cpu().numpy()
        return out

    def _convert(self, name, f_ref):
        ar, cur = self.arena, self.batch.cur
        A = _native.BricksConvertArgs()
        A.f, A.nbr, A.bidx = ar.f.data_ptr(), ar.nbr.data_ptr(), ar.bidx.data_ptr()
        A.lmask, A.smask, A.side = ar.lmask.data_ptr(), ar.smask.data_ptr(), ar.side.data_ptr()
        A.sstart, A.loff = ar.sstart.data_ptr(), ar.loff.data_ptr()
        A.slot_of_cell, A.n_cells = self.slot32.data_ptr(), self.N
        A.f_ref, A.parity = f_ref.data_ptr(), cur
        A.vel, A.rho = ar.mom.convert_ptrs(cur)
        A.storage, A.n_slots = STORAGES[self.batch.storage], ar.S
        _native.call(name, ctypes.byref(A), self.batch.stream_handle)

    def get_f(self):
        torch = _torch()
        self.batch.synchronize()
        with torch.cuda.stream(self.batch._tstream):
            f_ref = torch.empty((27, self.N), dtype=torch.float32, device=self.batch.dev)
            self._convert("cs_bricks_get_f", f_ref)
        self.batch.synchronize()
        return f_ref.cpu().numpy()

    def set_f(self, f):
        torch = _torch()
        self.batch.synchronize()
        with torch.cuda.stream(self.batch._tstream):
            f_ref = torch.from_numpy(np.ascontiguousarray(f, np.float32)).to(self.batch.dev)
            self._convert("cs_bricks_set_f", f_ref)
        self.batch.synchronize()

    def set_velocity(self, u):
        torch = _torch()
        f = self.get_f() if self.batch.storage == "fp16" else None
        self.batch.synchronize()
        with torch.cuda.stream(self.batch._tstream):
            uu = torch.from_numpy(np.ascontiguousarray(u, np.float32)).to(self.batch.dev)
            self.arena.mom.set_velocity(self.batch.cur, self.slot_dev, uu)
        self.batch.synchronize()
        if f is not None:
            self.set_f(f)

    def ok(self):
        """All densities of the case finite (device reduction)."""
        ar, b0, nb = self.arena, self.b0, int(self.lay["n_bricks"])
        x = ar.mom.flat(self.batch.cur, b0 * BC, (b0 + nb) * BC)
        _native.call("cs_reduce_finite", ctypes.c_void_p(x.data_ptr()), x.numel(),
                     ctypes.c_void_p(self.batch._work.data_ptr()), ctypes.c_void_p(self.batch._red.data_ptr()),
                     self.batch.stream_handle)
        self.batch.synchronize()
        return bool(self.batch._red[1].item() == 0)
