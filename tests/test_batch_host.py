"""Host-side pieces of the batching form: the range allocator and the arena footprint."""
import numpy as np

from cerebrosim_amd import batch as B


def test_ranges_first_fit_and_coalescing():
    r = B._Ranges(100)
    assert r.alloc(30) == 0 and r.alloc(20) == 30 and r.alloc(10) == 50
    assert r.high() == 60
    r.release(30, 20)
    assert r.alloc(25) == 60            # the 20-hole is too small: first fit goes past it
    assert r.alloc(15) == 30            # and the next smaller request takes the hole
    r.release(0, 30)
    r.release(30, 15)
    r.release(50, 10)
    r.release(60, 25)
    assert r.free == [(0, 100)] and r.high() == 0
    assert r.alloc(101) is None and r.alloc(0) == 0


def test_ranges_random_against_bitmap():
    rng = np.random.default_rng(3)
    r = B._Ranges(500)
    used = np.zeros(500, bool)
    live = []
    for _ in range(2000):
        if live and rng.random() < 0.45:
            s, n = live.pop(rng.integers(len(live)))
            assert used[s:s + n].all()
            used[s:s + n] = False
            r.release(s, n)
        else:
            n = int(rng.integers(1, 40))
            s = r.alloc(n)
            free_runs = [(i, j) for i, j in _runs(~used) if j - i >= n]
            if s is None:
                assert not free_runs
            else:
                assert s == free_runs[0][0] and not used[s:s + n].any()
                used[s:s + n] = True
                live.append((s, n))
        assert sum(ln for _, ln in r.free) == (~used).sum()
        hi = np.flatnonzero(used).max() + 1 if used.any() else 0
        assert r.high() == hi


def _runs(mask):
    out, i, n = [], 0, len(mask)
    while i < n:
        if mask[i]:
            j = i
            while j < n and mask[j]:
                j += 1
            out.append((i, j))
            i = j
        else:
            i += 1
    return out


def test_arena_bytes():
    fp32 = B.arena_bytes(10, 4, "fp32")
    fp16 = B.arena_bytes(10, 4, "fp16")
    assert fp32 - fp16 == 10 * (27 * 64 * 2 - 64 * 4)      # half the populations, one more density plane
    assert B.MAX_BRICKS * 27 * 64 < 2 ** 32


def test_packed_admission_arrays_round_trip():
    """pack_admission packs a layout's upload arrays (and a recorded sponge) into one buffer; each view of the buffer
    gives back the array in its device type and shape (here on a CPU tensor standing in for the device copy)."""
    import torch

    from cerebrosim_amd import bricks
    from cerebrosim_amd.geometry import open_pipe
    lat, _, _, _ = open_pipe((0, 0, 1), 10, lambda s: np.full_like(np.asarray(s, float), 4.0))
    lay = bricks.build(lat)
    rec = B.SetupRecorder(lat, 0.05, lay)
    rec.set_drive([0.01, 0.002, 0.0], period=40)
    rec.set_sponge(4, 0.7, kinds=("outlet",))
    packed = B.pack_admission(lay, rec)
    dev = torch.from_numpy(packed.buf.copy())
    got = lambda k: packed.view(dev, k).numpy()
    assert np.array_equal(got("slot"), lay["slot_of_cell"].astype(np.int32))
    assert np.array_equal(got("nbr"), lay["nbr"].reshape(-1).astype(np.int32))
    assert np.array_equal(got("lmask").view(np.uint32), lay["lmask"].reshape(-1))
    assert np.array_equal(got("qstart"), lay["qstart"][:-1].astype(np.int32))
    assert np.array_equal(got("qlink"), lay["qlink"][:int(lay["n_links"])].astype(np.float32))
    inl = np.flatnonzero((lay["lmask"] & np.uint32(1 << 27)).any(axis=1))
    assert np.array_equal(got("inl_idx"), inl)
    nr = int(lay["n_boundary"])
    assert np.array_equal(got("inl_rows"), lay["uin"].reshape(nr, -1)[inl].astype(np.float32))
    slots, own = rec.calls[-1][1]
    rows = np.flatnonzero(own)
    assert len(rows) and np.array_equal(got("om_rows"), rows)
    assert np.array_equal(got("om_data"), slots.reshape(-1, 64)[rows]) and np.array_equal(got("om_own"), own)
    assert all(o % 256 == 0 for o, _, _ in packed.offs.values())
