"""The simulate.py counterpart on the GPU: three compressed beats of a synthetic Y artery (0.1 mm, 5 us: the
reference's healthy grid) with the lumped circulation and the inlet calibration. The calibrated inflow reaches the
sourced 240 mL/min, the outlet resistances give the sourced 135 mL/min into the MCA branch, mass balances across the
three flow slabs, and the outputs have the reference's formats (simulate.py:356-371). The wall shear stress of the
stored frames (computed on the device at the frame cadence) equals the numpy restatement of wss.py's shear() applied
to the stored velocity frames, and its time averages follow their definitions (TAWSS = mean |tau|, OSI =
(1 - |mean tau| / mean |tau|) / 2, wss.py:120-122)."""
import json
import os

import numpy as np
import pytest

from oracle.wss import shear

pytestmark = pytest.mark.gpu


def test_simulate_y_artery(tmp_path):
    from cerebrosim_amd import simulate
    from cerebrosim_amd.geometry import artery_y
    lat, ends, probes = artery_y(0.1)
    ser, meta = simulate.run(lat, ends, probes, 1e-4, 5e-6, str(tmp_path), case="y_healthy", beats=3,
                             heart_period=0.04, log=lambda m: print(m, flush=True))
    n = len(ser["t"])
    assert n == 3 * 40 and ser["Q"].shape == (n, 3) and ser["p_open"].shape == (n, 3)
    assert ser["p_probe"].shape == (n, len(probes)) and ser["umax"].shape == (n,)
    last = ser["t"] > ser["t"][-1] - 0.04
    Q = ser["Q"][last].mean(0)
    print("mean flows of the last beat (mL/min):", Q, flush=True)
    assert abs(Q[0] / 240.0 - 1) < 0.03                       # calibrated inflow
    assert abs(Q[1] / 135.0 - 1) < 0.05                       # MCA share from the outlet resistances
    assert abs(Q[0] - Q[1] - Q[2]) < 0.02 * Q[0]              # mass balance across the slabs
    assert np.all(np.isfinite(ser["p_open"])) and 60 < ser["p_open"][last].mean() < 110
    f = np.load(os.path.join(tmp_path, "fields.npz"))
    N = lat["nbr"].shape[1]
    assert f["coords"].dtype == np.int16 and f["coords"].shape == (N, 3)          # 0.1 mm: one bin per cell
    assert f["p_mmhg"].dtype == np.float16 and f["p_mmhg"].shape == (40, N)
    assert f["u_m_s"].dtype == np.float16 and f["u_m_s"].shape == (40, N, 3)
    assert f["wall_mm"].dtype == np.float32 and f["t"].shape == (40,) and float(f["dx_mm"]) == pytest.approx(0.1)
    with open(os.path.join(tmp_path, "meta.json")) as fh:
        m = json.load(fh)
    assert m["steps_per_beat"] == 8000 and m["cells"] == N and m["done"]
    assert [e["label"] for e in m["open_ends"]] == ["inlet (ICA)", "MCA outlet", "ACA outlet"]
    assert all(p["cells"] > 0 for p in m["probes"])

    # wall shear stress over the stored frames
    from cerebrosim_amd import wss
    w = np.load(os.path.join(tmp_path, "wss.npz"))
    V = len(lat["wall_verts"])
    assert set(w.files) == {"verts", "tau", "tawss", "osi", "wall", "t"}
    assert w["tau"].shape == (40, V, 3) and w["tau"].dtype == np.float16 and w["wall"].all()
    cells = lat["origin"] + lat["coords"] * 0.1
    idx, c, _, _ = wss.weights(lat["wall_verts"], lat["wall_normals"], cells, lat["wall_mm"])
    errs = []
    for k in range(40):
        t_ref = shear(f["u_m_s"][k].astype(np.float64), lat["wall_normals"], idx, c.astype(np.float64), mu=wss.MU)
        t_dev = w["tau"][k].astype(np.float64)
        errs.append(np.linalg.norm(t_dev - t_ref) / np.linalg.norm(t_ref))
    print(f"WSS per frame vs the restatement on the stored frames: max rel L2 {max(errs):.2e}; TAWSS mean "
          f"{w['tawss'].mean():.3f} Pa", flush=True)
    assert max(errs) < 2e-3                                    # float16 frames and float16 tau
    tau = w["tau"].astype(np.float64)
    mag = np.linalg.norm(tau, axis=2)
    tawss = mag.mean(0)
    osi = 0.5 * (1 - np.linalg.norm(tau.mean(0), axis=1) / np.maximum(tawss, 1e-9))
    assert np.linalg.norm(w["tawss"] - tawss) / np.linalg.norm(tawss) < 2e-3
    assert np.abs(w["osi"] - osi).max() < 5e-3
    assert 0.5 < w["tawss"].mean() < 50 and (w["osi"] >= -1e-6).all() and (w["osi"] <= 0.5 + 1e-6).all()
