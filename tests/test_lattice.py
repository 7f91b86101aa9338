"""build_lattice against the reference's published cell and link counts (geometry/solver.json) and closed forms."""
import numpy as np
import pytest

from cerebrosim_amd import lattice as L
from cerebrosim_amd.geometry import open_pipe, periodic_pipe, poiseuille_box, stenosis_pipe


def test_tables():
    assert L.Q == 27 and L.C.shape == (27, 3)
    assert np.allclose(L.W.sum(), 1.0)
    assert (L.C[L.OPP] == -L.C).all()
    # links 1..6 are +x, -x, +y, -y, +z, -z (the finite-difference gradient relies on it)
    assert (L.C[1:7] == [[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]]).all()
    # isotropy up to 4th order (needed by the collision's moment identities)
    w = L.W
    c = L.C.astype(float)
    assert np.allclose((w[:, None, None] * c[:, :, None] * c[:, None, :]).sum(0), np.eye(3) / 3)
    m4 = np.einsum("i,ia,ib,ic,id->abcd", w, c, c, c, c)
    d = np.eye(3)
    iso = (np.einsum("ab,cd->abcd", d, d) + np.einsum("ac,bd->abcd", d, d) + np.einsum("ad,bc->abcd", d, d)) / 9
    assert np.allclose(m4, iso)


def test_oracle_tables_agree():
    import oracle
    c, w, o = oracle.tables()
    assert (c == L.C).all() and np.array_equal(w, L.W) and (o == L.OPP).all()


@pytest.mark.parametrize("name,axis,cells,wall,inlet,outlet", [
    ("straight", (0, 0, 1), 75600, 137890, 2818, 2818),        # geometry/solver.json:72-91
    ("tilted", (1, 2, 2), 75280, 136979, 2832, 2779),          # geometry/solver.json:93-112
])
def test_open_pipe_counts(name, axis, cells, wall, inlet, outlet):
    lat, sc, rc, ax = open_pipe(axis, 240, lambda s: np.full_like(s, 10.0, dtype=float))
    assert lat["nbr"].shape[1] == cells
    assert lat["counts"] == {"wall": wall, "inlet": inlet, "outlet": [outlet]}


@pytest.mark.parametrize("axis", [(0, 0, 1), (0, 1, 0), (1, 0, 0)])
def test_open_pipe_grid_axis_path_equals_general_form(monkeypatch, axis):
    """Pipes along a grid axis compute the signed distance on the grid's lines and planes: the lattice and the
    axial / radial coordinates have the same bits as the general (cells, 3) form."""
    import cerebrosim_amd.geometry as G

    def radius(s):
        s = np.asarray(s, float)
        return 9.3 * (1 - 0.35 * (1 - np.cos(2 * np.pi * np.clip((s - 20) / 30, 0, 1))))

    fast = G.open_pipe(axis, 90, radius)
    monkeypatch.setattr(G, "_grid_axis", lambda a: None)
    ref = G.open_pipe(axis, 90, radius)
    for k in ("nbr", "q", "uin", "coords"):
        assert np.array_equal(fast[0][k], ref[0][k]), k
    assert fast[0]["counts"] == ref[0]["counts"]
    assert np.array_equal(fast[1], ref[1]) and np.array_equal(fast[2], ref[2])


def test_stenosis_cells():
    lat, _, _, D = stenosis_pipe(1.0e-4)                  # geometry/solver.json:120 (20 cells across)
    assert lat["nbr"].shape[1] == 144678


@pytest.mark.slow
def test_stenosis_fine_cells():
    lat, _, _, D = stenosis_pipe(5.0e-5)                  # geometry/solver.json:162 (40 cells across)
    assert lat["nbr"].shape[1] == 1156558


def test_poiseuille_box_size():
    lat, r, R = poiseuille_box()
    assert lat["nbr"].shape[1] == 20160
    assert lat["counts"]["inlet"] == 0 and lat["counts"]["outlet"] == []
    assert (r < R).all()


def test_periodic_wrap_and_sources():
    lat, r, R = periodic_pipe(10)
    nbr, coords, shape = lat["nbr"], lat["coords"], np.array(lat["shape"])
    idx = {tuple(c): n for n, c in enumerate(coords)}
    for i in range(27):
        for n in range(0, len(coords), 7):
            src = tuple((coords[n] - L.C[i]) % [10**9, 10**9, shape[2]])
            if nbr[i, n] >= 0:
                assert idx[src] == nbr[i, n]
            else:
                assert nbr[i, n] == L.WALL and src not in idx


def test_wall_fraction_flat_wall():
    # flat wall at x = 2.3: phi = 2.3 - x < 0 for fluid x > 2.3; a link from x = 3 towards x = 2 crosses at 0.7
    nx, ny, nz = 8, 4, 4
    x = np.arange(nx)[:, None, None] * np.ones((1, ny, nz))
    lat = L.build_lattice(2.3 - x, [0, 0, 0], 1.0, periodic=(1, 2))
    coords = lat["coords"]
    n = np.flatnonzero((coords[:, 0] == 3))[0]
    # link 1 (+x) pulls from x - 1 = 2 (solid): q = -phi(3) / (phi(2) - phi(3)) = 0.7 / 1
    assert lat["nbr"][1, n] == L.WALL
    assert np.isclose(lat["q"][1, n], 0.7, atol=1e-6)
    # links that stay in the fluid keep q = 0.5 (unused) and sources in the grid
    assert lat["q"][2, n] == np.float32(0.5)
    # off-grid solid node (x = -1 from x = 0 is never fluid here); q of half-way default where phi is unknown
    assert np.isfinite(lat["q"]).all()


def test_off_grid_wall_gets_half():
    # fluid touching the grid edge: the solid source lies outside the grid -> q = 0.5 (lbm.py:311-315)
    phi = -np.ones((4, 4, 4), np.float32)
    lat = L.build_lattice(phi, [0, 0, 0], 1.0)
    miss = lat["nbr"] == L.WALL
    assert miss.any() and (lat["q"][miss] == np.float32(0.5)).all()


def test_inlet_profile_parabolic_unit_mean():
    lat, sc, rc, ax = open_pipe((0, 0, 1), 40, lambda s: np.full_like(s, 10.0, dtype=float))
    cells = np.flatnonzero((lat["nbr"] == L.INLET).any(0))
    uz = lat["uin"][cells, 2]
    o = lat["openings"][0]
    rel = lat["coords"][cells] - o["centre"]
    r = np.hypot(rel[:, 0], rel[:, 1])
    assert np.allclose(uz, np.clip(2 * (1 - (r / 10) ** 2), 0, None), atol=1e-6)   # inward = +z
    assert (lat["uin"][cells, :2] == 0).all()


def test_scatter_roundtrip():
    lat, r, R = periodic_pipe(10)
    g = L.scatter(lat, r)
    assert np.isnan(g).sum() == np.prod(lat["shape"]) - len(r)
    assert np.array_equal(g[tuple(lat["coords"].T)], r)


def test_native_builder_equals_numpy(monkeypatch):
    """The C++ builder (cs_lattice_build, the default) gives the numpy restatement's arrays and counts exactly, on
    periodic, tilted, narrowed and branching geometries with inlets and several outlets."""
    from cerebrosim_amd import geometry
    calls = []
    real = geometry.build_lattice

    def spy(phi, origin, dx, openings=(), periodic=(), opening_slack=1.0, **kw):
        calls.append((np.array(phi, np.float32), origin, dx, list(openings), tuple(periodic), opening_slack))
        return real(phi, origin, dx, openings, periodic, opening_slack)

    monkeypatch.setattr(geometry, "build_lattice", spy)
    geometry.periodic_pipe(20, nz=4)
    geometry.poiseuille_box()
    geometry.open_pipe((1, 2, 2), 30, lambda s: np.full_like(np.asarray(s, float), 5.0))
    geometry.stenosis_pipe(1.5e-4, sev=0.6)
    geometry.bifurcation(0.12)
    assert len(calls) == 5
    for phi, origin, dx, ops, per, slack in calls:
        a = L.build_lattice(phi, origin, dx, ops, per, slack, native=True, threads=3)
        b = L.build_lattice_numpy(phi, origin, dx, ops, per, slack)
        for k in ("coords", "nbr", "q", "uin"):
            assert a[k].dtype == b[k].dtype and np.array_equal(a[k], b[k]), (phi.shape, k)
        assert a["counts"] == b["counts"] and a["n_outlets"] == b["n_outlets"]
