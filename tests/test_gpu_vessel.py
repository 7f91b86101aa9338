"""The story artery on the GPU against the reference's published results.

* ``cs_sdf_tubes`` against the numpy restatement of k_sdf (oracle/sdf.py) on sampled voxels (random, and within
  0.15 mm of the wall where the sign decides the lattice).
* The lattice built from it (validate.py:294-308) against geometry/solver.json "artery": fluid cells and every link
  count, exactly.
* Test 5 of validate.py (validate.py:311-340): the maximum speed after 21000 steps against the published value.
* The simulate.py runs of the healthy case (4 heartbeats at 0.1 mm / 5 us, inlet calibrated on the run) and of the
  occluded case (3 heartbeats on the healthy calibration) against reports/step4_report.md:7-21: mean inflow, MCA and
  ACA flows, peak speed and the probe pressures over the last beat, and the largest speed in the stored frames
  against reports/step5_report.md:40-44.
* The healthy run again with FP16 population storage (the precision form on this caller) against the FP32 run.
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
SEGMENT = os.path.join(GOLDEN, "segment_story.json")


@pytest.fixture(scope="module")
def ref():
    with open(os.path.join(GOLDEN, "reference_data.json")) as fh:
        return json.load(fh)


def test_sdf_kernel_matches_the_oracle():
    from cerebrosim_amd import vessel as V
    from oracle.sdf import sdf_points
    g = V.build("healthy", segment=SEGMENT)
    phi, lo = V.sdf_grid(g, 0.1)
    assert phi.shape == V.grid_shape(*V.bounds(g), 0.1)
    flat = phi.ravel()
    rng = np.random.default_rng(0)
    near = np.flatnonzero(np.abs(flat) < 0.15)
    pick = np.concatenate([rng.integers(0, flat.size, 20000), rng.choice(near, 20000, replace=False)])
    P = lo + np.stack(np.unravel_index(pick, phi.shape), 1) * 0.1
    d = sdf_points(P, *V.packed(g), V.JUNCTION_BLEND)
    err = np.abs(flat[pick] - d).max()
    print(f"cs_sdf_tubes vs oracle: max |diff| {err:.2e} mm over {len(pick)} voxels", flush=True)
    assert err < 5e-5
    assert ((flat[pick] < 0) == (d < 0)).all()


def test_artery_lattice_and_speed_match_the_published_run(ref):
    from cerebrosim_amd import validate
    res, beat = validate.artery(SEGMENT)
    pub = ref["artery"]
    print(f"artery: {res['cells']} cells, {res['counts']}, {res['mlups']:.0f} MLUPS, u_max {res['u_max_m_s']:.7f} "
          f"(published {pub['u_max_m_s']:.7f})", flush=True)
    assert res["cells"] == pub["cells"] and res["counts"] == pub["counts"]
    assert res["steps"] == pub["steps"] and res["ok"]
    assert abs(res["u_max_m_s"] / pub["u_max_m_s"] - 1) < 1e-5
    assert beat["healthy"]["steps_per_beat"] == pytest.approx(200000)


def _run_case(case, tmp, cal=None, storage="fp32"):
    from cerebrosim_amd import simulate
    from cerebrosim_amd import vessel as V
    lat, ends, probes = V.artery_case(case, segment=SEGMENT)
    dx, dt = V.GRIDS[case]
    beats = {"healthy": 4, "occluded": 3}[case]                  # the report's runs (step4_report.md:7-9)
    ser, meta = simulate.run(lat, ends, probes, dx, dt, str(tmp), case=case, beats=beats, cal=cal,
                             log=lambda m: print(m, flush=True), storage=storage)
    return lat, probes, ser, meta, tmp


@pytest.fixture(scope="module")
def healthy_run(tmp_path_factory):
    return _run_case("healthy", tmp_path_factory.mktemp("healthy"))


def _check_fields(out_dir, pub, tol=0.02):
    """reports/step5_report.md:40-44: the largest speed in the stored frames (fields.npz, 40 frames of the last
    beat on 0.1 mm bins, float16), the figure the reference's renderer reads."""
    f = np.load(os.path.join(str(out_dir), "fields.npz"))
    assert f["u_m_s"].shape[0] == 40 and f["u_m_s"].dtype == np.float16
    vmax = float(np.linalg.norm(f["u_m_s"].astype(np.float32), axis=2).max())
    print(f"stored fields: max speed {vmax:.3f} m/s (published {pub['fields_max_speed_m_s']})", flush=True)
    assert abs(vmax - pub["fields_max_speed_m_s"]) < tol


def _check_last_beat(ser, probes, pub, tol_q=1.0, tol_p=0.15, tol_u=0.01):
    last = ser["t"] > ser["t"][-1] - 1.0
    Q = ser["Q"][last].mean(0)
    print(f"last beat: inflow {Q[0]:.2f}, MCA {Q[1]:.2f}, ACA {Q[2]:.2f} mL/min (published {pub['inflow_mL_min']}, "
          f"{pub['mca_mL_min']}, {pub['aca_mL_min']}); peak speed {ser['umax'][last].max():.3f} m/s (published "
          f"{pub['peak_speed_m_s']})", flush=True)
    assert abs(Q[0] - pub["inflow_mL_min"]) < tol_q, Q
    assert abs(Q[1] - pub["mca_mL_min"]) < tol_q and abs(Q[2] - pub["aca_mL_min"]) < tol_q, Q
    assert abs(ser["umax"][last].max() - pub["peak_speed_m_s"]) < tol_u
    p = ser["p_probe"][last]
    for k, pr in enumerate(probes):
        want = pub["probe_mmHg"][pr["name"]]
        if want is None:                                           # no lumen at the probe
            assert pr["cells"] == 0 and np.isnan(p[:, k]).all(), pr["name"]
            continue
        v = float(np.mean(p[:, k]))
        print(f"{pr['name']}: {v:.2f} mmHg (published {want})", flush=True)
        assert abs(v - want) < tol_p, pr["name"]


def test_healthy_heartbeats_match_the_step4_report(ref, healthy_run):
    lat, probes, ser, meta, out = healthy_run
    assert lat["nbr"].shape[1] == ref["artery"]["cells"] and meta["cells"] == ref["artery"]["cells"]
    _check_last_beat(ser, probes, ref["step4_healthy"])
    _check_fields(out, ref["step4_healthy"])


def test_occluded_heartbeats_match_the_step4_report(ref, healthy_run, tmp_path):
    """The full occlusion with the healthy run's calibrated inlet pressure and resistances (simulate.py:323-332):
    the lumen splits at the clot, the MCA carries nothing and the ACA takes the inflow."""
    lat, probes, ser, meta, out = _run_case("occluded", tmp_path, cal=healthy_run[3]["calibration"])
    _check_last_beat(ser, probes, ref["step4_occluded"])
    _check_fields(out, ref["step4_occluded"])


def test_healthy_heartbeats_in_fp16_storage(healthy_run, tmp_path):
    """The precision form on the simulate.py caller: the same 4 calibrated heartbeats with the populations stored
    as FP16 equilibrium deviations (FP32 arithmetic), against the FP32 run. Series relative L2 over the last beat."""
    _, probes, ser32, meta32, _ = healthy_run
    _, _, ser16, meta16, _ = _run_case("healthy", tmp_path, storage="fp16")
    last = ser32["t"] > ser32["t"][-1] - 1.0
    assert np.array_equal(ser16["t"], ser32["t"])

    def rel(a, b):
        return float(np.sqrt(np.nansum((a - b) ** 2) / np.nansum(b ** 2)))

    errs = {"Q": rel(ser16["Q"][last], ser32["Q"][last]),
            "p_probe": rel(ser16["p_probe"][last], ser32["p_probe"][last]),
            "umax": rel(ser16["umax"][last], ser32["umax"][last])}
    print(f"FP16 vs FP32 storage, last heartbeat, relative L2: {errs}; inlet calibration R_seg "
          f"{meta16['calibration']['R_seg']:.4g} vs {meta32['calibration']['R_seg']:.4g}", flush=True)
    assert errs["Q"] < 1e-3 and errs["p_probe"] < 1e-3 and errs["umax"] < 1e-3, errs
