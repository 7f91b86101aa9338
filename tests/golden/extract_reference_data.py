"""Copies the numbers the tests compare against out of the reference's committed data files (read as JSON/text, no
reference code is run) into tests/golden/reference_data.json:

* geometry/physics.json: the inflow waveform over one beat, 200 samples rounded to 5 decimals (vessel.py:253-265
  generated it; physics.json:"inlet");
* reports/step4_report.md:28: the calibrated circuit's resistances (GPa s/m^3);
* geometry/solver.json:"artery": the healthy artery at 0.1 mm (validate.py:311-340): cells, link counts, openings,
  steps, speed and the maximum speed after them;
* reports/step4_report.md:7-21: each case's grid, beats, mean inflow, MCA and ACA flows, peak speed and probe
  pressures (None where the report says there is no lumen);
* reports/step5_report.md:40-44: each case's maximum speed in the stored fields (40 frames of the last beat, 0.1 mm
  bins).

    python tests/golden/extract_reference_data.py [/path/to/reference]
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def main(ref):
    phys = json.load(open(os.path.join(ref, "geometry", "physics.json")))
    inlet = phys["inlet"]
    report = open(os.path.join(ref, "reports", "step4_report.md")).read()
    m = re.search(r"outlet resistances MCA ([\d.]+), ACA ([\d.]+) GPa s/m\^3; segment \(incl. extensions\) ([\d.]+)",
                  report)
    out = {"inflow_waveform": {"t_s": inlet["waveform_t_s"], "q_over_mean": inlet["waveform_Q_over_mean"],
                               "mean_flow_mL_min": inlet["mean_flow_mL_min"],
                               "pulsatility_index": inlet["pulsatility_index"],
                               "source": "geometry/physics.json: inlet"},
           "circuit_GPa_s_m3": {"R_mca": float(m.group(1)), "R_aca": float(m.group(2)), "R_seg": float(m.group(3)),
                                "source": "reports/step4_report.md:28"}}
    art = json.load(open(os.path.join(ref, "geometry", "solver.json")))["artery"]
    out["artery"] = {k: art[k] for k in ("cells", "steps", "secs", "mlups", "u_max_m_s", "counts", "openings",
                                         "U_mean_inlet")}
    out["artery"]["source"] = "geometry/solver.json: artery"
    cases = ("healthy", "narrow70", "occluded")
    probes = {c: {} for c in cases}
    for name in ("ICA", "junction", "MCA before clot", "clot throat", "MCA after clot", "MCA distal", "ACA distal"):
        cells = re.search(r"\| " + name + r" \| ([^|]+) \| ([^|]+) \| ([^|]+) \|", report).groups()
        for c, v in zip(cases, cells):
            v = v.strip()
            probes[c][name] = None if v.startswith("(") else float(v)
    for c in cases:
        row = re.search(r"\| " + c + r" \| ([\d.]+) mm \| (\d+) \| (\d+) mL/min \| (-?\d+) \| (-?\d+) \| [^|]+ \| "
                        r"([\d.]+) m/s", report)
        out["step4_" + c] = {"dx_mm": float(row.group(1)), "beats": int(row.group(2)),
                             "inflow_mL_min": float(row.group(3)), "mca_mL_min": float(row.group(4)),
                             "aca_mL_min": float(row.group(5)), "peak_speed_m_s": float(row.group(6)),
                             "probe_mmHg": probes[c], "source": "reports/step4_report.md:7-21"}
    step5 = open(os.path.join(ref, "reports", "step5_report.md")).read()
    for c in cases:
        row = re.search(r"\| " + c + r" \| ([\d.]+) m/s \| ([\d.]+) m/s \| ([\d.]+) m/s \|", step5)
        out["step4_" + c]["fields_max_speed_m_s"] = float(row.group(2))
    json.dump(out, open(os.path.join(HERE, "reference_data.json"), "w"), indent=1)
    print("wrote reference_data.json")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
