"""Fits the 10-harmonic trigonometric series of the reference's published inflow waveform (geometry/physics.json
"inlet": Q(t) / Q_mean at 200 times over one beat, 5 decimals; copied into reference_data.json) by least squares,
and prints the coefficients [a0, a1, b1, ..., a10, b10] of a0 + sum_k a_k cos(2 pi k t / T) + b_k sin(2 pi k t / T)
that cerebrosim_amd/artery.py holds as INFLOW_HARMONICS.

    python tests/golden/fit_inflow_harmonics.py
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def fit(t_over_T, y, nh=10):
    cols = [np.ones_like(t_over_T)]
    for k in range(1, nh + 1):
        cols += [np.cos(2 * np.pi * k * t_over_T), np.sin(2 * np.pi * k * t_over_T)]
    coef, *_ = np.linalg.lstsq(np.stack(cols, 1), y, rcond=None)
    return coef


def main():
    w = json.load(open(os.path.join(HERE, "reference_data.json")))["inflow_waveform"]
    t = np.asarray(w["t_s"], float)
    y = np.asarray(w["q_over_mean"], float)
    coef = fit(t / (t[1] - t[0]) / len(t), y)
    print("INFLOW_HARMONICS = (" + ", ".join(f"{c:.9f}" for c in coef) + ")")


if __name__ == "__main__":
    main()
