"""Brick layout on the CPU: builder invariants, and the AA-pattern design (parity forms, boundary redirects, side
copies, conversions) checked against the FP64 oracle through a numpy model of the stepping (tests/aa_model.py)."""
import numpy as np
import pytest

import oracle
from aa_model import AAModel
from cerebrosim_amd import bricks
from cerebrosim_amd import lattice as L
from cerebrosim_amd.geometry import open_pipe, periodic_pipe, poiseuille_box


def tilted_pipe():
    lat, _, _, _ = open_pipe((1, 2, 2), 10, lambda s: np.full_like(np.asarray(s, float), 3.0))
    return lat


def test_builder_invariants():
    lat = tilted_pipe()
    lay = bricks.build(lat)
    N = lat["nbr"].shape[1]
    slot = lay["slot_of_cell"]
    assert len(np.unique(slot)) == N
    assert (lay["cell_of_slot"][slot] == np.arange(N)).all()
    assert (lay["cell_of_slot"] >= 0).sum() == N
    # brick coordinates and local positions reproduce the cell coordinates
    b, l = slot // 64, slot % 64
    pos = lay["coords"][b] * 4 + np.stack([l % 4, (l // 4) % 4, l // 16], 1)
    assert (pos == lat["coords"]).all()
    # every fluid link is the implicit neighbour through the brick table
    for i in range(27):
        m = lat["nbr"][i] >= 0
        p = lat["coords"][m] - L.C[i]
        bp = np.floor_divide(p, 4)
        o = bp - lay["coords"][b[m]]
        d = (o[:, 0] + 1) + 3 * (o[:, 1] + 1) + 9 * (o[:, 2] + 1)
        nbk = lay["nbr"][b[m], d]
        lp = p - bp * 4
        assert (nbk * 64 + lp[:, 0] + 4 * lp[:, 1] + 16 * lp[:, 2] == slot[lat["nbr"][i][m]]).all()
    # boundary masks agree with the neighbour table
    bidx = lay["bidx"][b]
    bm = bidx >= 0
    lm = lay["lmask"][bidx[bm], l[bm]]
    for i in range(27):
        assert (((lm >> i) & 1).astype(bool) == (lat["nbr"][i][bm] >= 0)).all()
    inl = (lat["nbr"] == L.INLET).any(0)
    assert ((((lm >> 27) & 1).astype(bool)) == inl[bm]).all()
    assert not inl[~bm].any()
    # interior bricks: every cell fluid, every link fluid
    for bk in lay["interior"]:
        cells = lay["cell_of_slot"][bk * 64:(bk + 1) * 64]
        assert (cells >= 0).all() and (lat["nbr"][:, cells] >= 0).all()


def test_periodic_box_wraps():
    lat, r, R = poiseuille_box()
    lay = bricks.build(lat)
    assert lay["periodic"] == [0, 0, 1]
    assert lay["n_cells"] == 20160


def test_periodic_inference_matches_brute_force():
    """The face-cell test of infer_periodic equals the definition: some fluid link whose source coordinate differs
    from coords - c by a wrap."""
    def brute(lat):
        nbr, coords = lat["nbr"], lat["coords"].astype(np.int64)
        per = [0, 0, 0]
        for i in range(1, 27):
            m = nbr[i] >= 0
            d = coords[nbr[i][m]] - coords[m] + L.C[i]
            for a in range(3):
                per[a] |= int(np.any(d[:, a] != 0))
        return per
    lats = [periodic_pipe(8, nz=4)[0], poiseuille_box()[0], tilted_pipe(),
            open_pipe((0, 0, 1), 12, lambda s: np.full_like(np.asarray(s, float), 3.0))[0]]
    for lat in lats:
        assert bricks.infer_periodic(lat) == brute(lat)
    assert bricks.infer_periodic(lats[0]) == [0, 0, 1] and bricks.infer_periodic(lats[2]) == [0, 0, 0]


def test_short_periodic_axis_is_rejected():
    lat, _, _ = periodic_pipe(10)           # 2 nodes along the periodic axis: not a multiple of the brick edge
    with pytest.raises(bricks.LayoutError):
        bricks.build(lat)


def test_mismatched_table_is_rejected():
    lat = tilted_pipe()
    nbr = lat["nbr"].copy()
    n = np.flatnonzero(nbr[1] >= 0)[0]
    nbr[1, n] = nbr[1, np.flatnonzero(nbr[1] >= 0)[5]]
    bad = dict(lat)
    bad["nbr"] = nbr
    with pytest.raises(bricks.LayoutError):
        bricks.build(bad)


@pytest.mark.parametrize("par", [0, 1])
def test_conversion_roundtrip(par):
    lat = tilted_pipe()
    lay = bricks.build(lat)
    m = AAModel(lat, lay, 0.05)
    f = np.random.default_rng(1).standard_normal((27, m.N))
    m.set_f(f, par)
    assert np.array_equal(m.get_f(par), f)


@pytest.mark.parametrize("start_par", [0, 1])
def test_aa_stepping_matches_oracle(start_par):
    lat = tilted_pipe()
    lay = bricks.build(lat)
    nu, force, p_out = 0.05, (1e-5, 2e-5, -1e-5), [0.001]
    rng = np.random.default_rng(7)
    f0 = 1e-3 * rng.standard_normal((27, lat["nbr"].shape[1]))
    u0 = 0.02 * rng.standard_normal((lat["nbr"].shape[1], 3))
    o = oracle.OracleSolver(lat, nu, model="regularised", dtype=np.float64, threads=1)
    o.set_force(force)
    o.set_outlet_pressure(p_out)
    o.f[o.cur][:] = f0
    o.vels[o.cur][:] = u0
    m = AAModel(lat, lay, nu, force=np.asarray(force, np.float32).astype(float), p_out=p_out)
    m.set_f(f0, start_par)
    m.vel = u0.copy()
    par = start_par
    for step in range(6):
        o.step_once(1.0)
        m.step(par)
        par = 1 - par
        ref = o.get_f()
        got = m.get_f(par)
        assert np.abs(got - ref).max() < 1e-13 * max(1.0, np.abs(ref).max()) * 100, step
        assert np.allclose(m.vel, o.vels[o.cur], atol=1e-15)


def test_slot_table_groups():
    """The slot table is a permutation of the 1728 block offsets; the odd-step accesses of each neighbour brick form
    one contiguous record (1000 own slots, 100 per face, 10 per edge, 1 per corner)."""
    t = bricks.slot_table().astype(np.int64)
    assert sorted(t.ravel().tolist()) == list(range(27 * 64))
    l = np.arange(64)
    pos = np.stack([l % 4, (l // 4) % 4, l // 16], 1)
    groups = {}
    for k in range(27):
        src = pos - L.C[k]
        o = np.where(src < 0, 0, np.where(src > 3, 2, 1))
        g = o[:, 0] + 3 * o[:, 1] + 9 * o[:, 2]
        for gg, off in zip(g, t[k]):
            groups.setdefault(int(gg), []).append(int(off))
    sizes = {g: len(v) for g, v in groups.items()}
    assert sizes[13] == 1000
    for g, v in groups.items():
        v = sorted(v)
        assert v == list(range(v[0], v[0] + len(v))), g          # contiguous
        nz = sum(1 for a in (g % 3, (g // 3) % 3, g // 9) if a != 1)
        assert len(v) == {0: 1000, 1: 100, 2: 10, 3: 1}[nz]


def test_odd_step_staging_maps_to_direct_slots():
    """Interior bricks stage their odd-step footprint (own group + one record from each of the 26 neighbour blocks)
    in LDS. Every staged position must stand for exactly the global slot the direct path accesses."""
    slots = bricks.slot_table().astype(np.int64)
    stage = bricks.stage_table().astype(np.int64)
    # staged order: own group, then the face, edge and corner neighbours, each in offset order
    nonown = lambda o: sum(d != 1 for d in (o % 3, (o // 3) % 3, o // 9))
    order = [13] + sorted((o for o in range(27) if o != 13), key=lambda o: (nonown(o), o))
    # group g of a block: its slots, in block order
    l = np.arange(64)
    pos = np.stack([l % 4, (l // 4) % 4, l // 16], 1)
    members = {}
    for k in range(27):
        src = pos - L.C[k]
        o = np.where(src < 0, 0, np.where(src > 3, 2, 1))
        g = o[:, 0] + 3 * o[:, 1] + 9 * o[:, 2]
        for gg, off in zip(g, slots[k]):
            members.setdefault(int(gg), []).append(int(off))
    start = {g: min(v) for g, v in members.items()}
    # staged footprint: (neighbour offset o, block offset)
    foot = []
    for o in order:
        g = 13 if o == 13 else 26 - o
        foot += [(o, start[g] + j) for j in range(len(members[g]))]
    assert len(foot) == 27 * 64
    for i in range(27):
        src = pos - L.C[i]
        o = np.where(src < 0, 0, np.where(src > 3, 2, 1))
        oi = o[:, 0] + 3 * o[:, 1] + 9 * o[:, 2]
        ml = (src[:, 0] % 4) + 4 * (src[:, 1] % 4) + 16 * (src[:, 2] % 4)
        direct = [(int(a), int(slots[L.OPP[i]][m])) for a, m in zip(oi, ml)]
        staged = [foot[p] for p in stage[i]]
        assert staged == direct, i
    # a bijection: the 27 x 64 reads cover the footprint once
    assert sorted(stage.ravel().tolist()) == list(range(27 * 64))
    # face and edge records start at even block offsets and even staged positions (float16 moves them as 4-byte
    # pairs); the records of the neighbours below (oz = -1) come before those of oz = 0, then oz = +1
    for g, v in members.items():
        if len(v) in (100, 10):
            assert start[g] % 2 == 0 and sorted(v) == list(range(start[g], start[g] + len(v))), g
    by_plane = sorted((g for g in members if g != 13), key=lambda g: start[g])
    assert [g // 9 for g in by_plane] == sorted(g // 9 for g in by_plane)
    pos_staged = 1000
    for o in order[1:]:
        size = len(members[26 - o])
        assert size == 1 or pos_staged % 2 == 0, o
        pos_staged += size
    # the packed record table the kernels load the footprint with
    rec = bricks.record_table().astype(np.int64)
    assert foot[:1000] == [(13, j) for j in range(1000)]
    assert [(int(e >> 11), int(e & 2047)) for e in rec] == foot[1000:]


def test_compact_link_arrays():
    """qlink / side order: brick by brick, lane by lane, direction 1..26; qstart / sstart the brick runs, loff the lane
    runs, ldir the per-brick unions; and the kernels' index (lane offset + popcount of the lane's links below i) finds
    every entry. Checked against the dense qc, lmask and smask with plain loops."""
    lat = tilted_pipe()
    lay = bricks.build(lat, dense=True)
    nb = lay["n_boundary"]
    qc = lay["qc"].reshape(nb, 27, 64)
    # the native builder's arrays equal the numpy restatement's
    ref = bricks.compact_links(lay)
    for k, v in ref.items():
        assert np.array_equal(np.asarray(lay[k]), np.asarray(v)) and np.asarray(lay[k]).dtype == np.asarray(v).dtype \
            if isinstance(v, np.ndarray) else lay[k] == v, k
    expect, sides = [], []
    starts, sstarts = [0], [0]
    for bb in range(nb):
        wq = ws = 0
        for l in range(64):
            m, sm = int(lay["lmask"][bb, l]), int(lay["smask"][bb, l])
            off = int(lay["loff"][bb, l])
            assert (off & 0xFFFF) == len(expect) - starts[-1] and (off >> 16) == len(sides) - sstarts[-1]
            nq = ns = 0
            for i in range(1, 27):
                if m and not (m >> i) & 1:
                    assert (off & 0xFFFF) + nq == len(expect) - starts[-1]
                    expect.append(qc[bb, i, l])
                    nq += 1
                    wq |= 1 << i
                if (sm >> i) & 1:
                    assert (off >> 16) + ns == len(sides) - sstarts[-1]
                    sides.append((bb, i, l))
                    ns += 1
                    ws |= 1 << i
        assert lay["ldir"][bb].tolist() == [wq, ws]
        starts.append(len(expect))
        sstarts.append(len(sides))
    assert np.array_equal(lay["qlink"][:-1], np.asarray(expect, np.float32)) and lay["qlink"][-1] == 0.0
    assert lay["qstart"].tolist() == starts and lay["sstart"].tolist() == sstarts
    assert lay["n_links"] == len(expect) and lay["n_side"] == len(sides)
    # every side link is a missing wall link with q >= 1/2
    for bb, i, l in sides:
        assert not (int(lay["lmask"][bb, l]) >> i) & 1 and qc[bb, i, l] >= 0.5


def run_fp16_model(steps, bounced_dir=True):
    """FP16-form AA model and FP64 oracle on the tilted pipe from rest, inlet drive 0.03, outlet p = 0."""
    lat = tilted_pipe()
    lay = bricks.build(lat)
    nu = 0.05
    o = oracle.OracleSolver(lat, nu, model="regularised", dtype=np.float64, threads=1)
    o.set_outlet_pressure([0.0])
    m = AAModel(lat, lay, nu, p_out=[0.0], storage="fp16")
    m.bounced_dir = bounced_dir
    par = 0
    for _ in range(steps):
        o.step_once(0.03)
        m.step(par, drive=0.03)
        par = 1 - par
    uo = o.vels[o.cur]
    return float(np.linalg.norm(m.vel - uo) / np.linalg.norm(uo))


def test_aa_fp16_form_matches_fp64_oracle():
    """The FP16 equilibrium-deviation storage rule (writer's moments, bounced links decoded with their own
    direction) keeps u within 1e-3 of the FP64 oracle; decoding bounced populations with the link's direction (a
    planted mistake) does not."""
    e = run_fp16_model(120)
    assert e <= 1e-3, e
    bad = run_fp16_model(120, bounced_dir=False)
    assert bad > 10 * max(e, 1e-4), (e, bad)


def test_census_counts_bricks_and_boundary_cells():
    """bricks.census on a hand-checkable lattice: a fully periodic 8^3 box is all interior at edge 4 and 8; a 6^3
    non-periodic box has every cell in a boundary brick (missing links at the faces) and 8 bricks of 4^3."""
    from cerebrosim_amd import bricks
    from cerebrosim_amd import lattice as L
    box = L.build_lattice(-np.ones((8, 8, 8), np.float32), [0, 0, 0], 1.0, periodic=(0, 1, 2))
    for edge in (4, 8):
        c = bricks.census(box, edge)
        assert c["bricks"] == (8 // edge) ** 3 and c["interior_bricks"] == c["bricks"]
        assert c["slots_per_cell"] == 1.0 and c["boundary_cell_fraction"] == 0.0
    small = L.build_lattice(-np.ones((6, 6, 6), np.float32), [0, 0, 0], 1.0)
    c = bricks.census(small, 4)
    assert c["bricks"] == 8 and c["interior_bricks"] == 0 and c["boundary_cell_fraction"] == 1.0
    assert c["slots"] == 512 and abs(c["slots_per_cell"] - 512 / 216) < 1e-12
