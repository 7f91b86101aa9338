"""Host-side pieces of the simulate.py counterpart against the reference's committed data and closed forms."""
import json
import math
import os

import numpy as np

from cerebrosim_amd import artery

HERE = os.path.dirname(os.path.abspath(__file__))
REF = json.load(open(os.path.join(HERE, "golden", "reference_data.json")))


def test_inflow_waveform_matches_physics_json():
    """The harmonics reproduce the published samples (5 decimals) to their rounding, with mean 1 and the sourced
    pulsatility (max - min) / mean."""
    t, q = artery.inflow_waveform(200)
    w = REF["inflow_waveform"]
    assert np.allclose(t, np.asarray(w["t_s"]), rtol=0, atol=1e-12)
    assert np.abs(q - np.asarray(w["q_over_mean"])).max() < 6e-6
    assert abs(q.mean() - 1) < 1e-12
    assert abs((q.max() - q.min()) - artery.PULSATILITY) < 1e-5      # on the 200 samples, as published


def test_circuit_matches_published_resistances():
    c = artery.circuit(REF["circuit_GPa_s_m3"]["R_seg"] * 1e9)
    assert round(c["R_out"][1] / 1e9, 2) == REF["circuit_GPa_s_m3"]["R_mca"]
    assert round(c["R_out"][2] / 1e9, 2) == REF["circuit_GPa_s_m3"]["R_aca"]
    # mean inlet pressure = (R_eq + R_seg) Q_in: the sourced inflow through segment + outlets in parallel
    q_in = artery.Q_INLET_MEAN * 1e-6 / 60
    assert math.isclose(c["P_in_coef"][0], (c["R_eq"] + c["R_seg"]) * q_in, rel_tol=1e-12)
    # outlets in parallel give the equivalent resistance p_out / q_in
    r1, r2 = c["R_out"][1], c["R_out"][2]
    assert math.isclose(1 / (1 / r1 + 1 / r2), c["R_eq"], rel_tol=1e-12)


def test_series_basis_and_orthogonality():
    """series() is a0 + sum a_k cos(2 pi k x) + b_k sin(2 pi k x): single terms give the plain cosines and sines,
    and the discrete means over a period recover each coefficient (orthogonality of the basis)."""
    x = np.arange(64) / 64
    assert np.allclose(artery.series([0.0, 0.0, 0.0, 1.0, 0.0], x), np.cos(4 * np.pi * x), atol=1e-14)
    assert np.allclose(artery.series([0.0, 0.0, 1.0], x), np.sin(2 * np.pi * x), atol=1e-14)
    c0 = np.array([1.0, 0.3, -0.2, 0.05, 0.1, 0.0, -0.02])
    y = artery.series(c0, x)
    assert abs(y.mean() - c0[0]) < 1e-14
    for k in (1, 2, 3):
        assert abs(2 * (y * np.cos(2 * np.pi * k * x)).mean() - c0[2 * k - 1]) < 1e-14
        assert abs(2 * (y * np.sin(2 * np.pi * k * x)).mean() - c0[2 * k]) < 1e-14


def test_step_size_whole_steps_per_sample():
    for dt, scale in ((5e-6, 1.0), (5e-6, 0.75), (2e-6, 0.6), (1e-6, 1.0)):
        d = artery.step_size(1e-4, dt, scale)
        n = artery.SAMPLE_S / d
        assert abs(n - round(n)) < 1e-9
        assert abs(d - dt * scale) / (dt * scale) < 0.01


def test_flux_slabs_select_cells_inside_each_end():
    from cerebrosim_amd.geometry import open_pipe
    lat, sc, rc, ax = open_pipe((0, 0, 1), 40, lambda s: np.full_like(np.asarray(s, float), 5.0))
    ends = lat["openings"]
    fl = artery.flux_slabs(lat, ends, 1.0, slab=(3.0, 5.0))
    assert all(a.flags["C_CONTIGUOUS"] for a in fl.values())    # passed to C as flat buffers
    for k, e in enumerate(ends):
        cells = fl["idx"][fl["slot"] == k]
        depth = -((lat["coords"][cells] - e["centre"]) @ np.asarray(e["normal"]))
        assert cells.size and (depth >= 3).all() and (depth < 5).all()
        w = fl["w"][fl["slot"] == k]
        sign = -1 if k == 0 else 1
        assert np.allclose(w, sign * np.asarray(e["normal"]) / 2.0)
