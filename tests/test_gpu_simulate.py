"""The simulate.py counterpart on the GPU: three compressed beats of a synthetic Y artery (0.1 mm, 5 us: the
reference's healthy grid) with the lumped circulation and the inlet calibration. The calibrated inflow reaches the
sourced 240 mL/min, the outlet resistances give the sourced 135 mL/min into the MCA branch, mass balances across the
three flow slabs, and the outputs have the reference's formats (simulate.py:356-371)."""
import json
import os

import numpy as np
import pytest

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
