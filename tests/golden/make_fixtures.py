"""Writes tests/golden/pipe_small.npz: frozen FP32 oracle results on a small tilted pipe with wall, inlet and outlet
links, from a seeded random state. The reference itself cannot run here (its kernels need NVIDIA Warp, which is not
installed), so these vectors come from the oracle; they freeze its output so that any later change to the oracle or
the GPU path that moves results is caught.

    python tests/golden/make_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from cerebrosim_amd.geometry import open_pipe  # noqa: E402

CASES = {
    # name: (model, sigma, bouzidi, steps recorded)
    "reg": ("regularised", 1.0, True, (1, 2, 50)),
    "trt": ("trt", 1.0, True, (50,)),
    "hyb": ("regularised", 0.8, True, (50,)),
    "halfway": ("regularised", 1.0, False, (50,)),
}
NU = 0.05
DRIVE = ([0.02, 0.0, 0.01], 40, 10)
FORCE = (1e-5, 2e-5, -1e-5)
P_OUT = [0.001]


def lattice():
    lat, _, _, _ = open_pipe((1, 2, 2), 10, lambda s: np.full_like(np.asarray(s, float), 3.0))
    return lat


def initial_state(N, seed=2024):
    rng = np.random.default_rng(seed)
    f = (1e-3 * rng.standard_normal((27, N))).astype(np.float32)
    u = (0.02 * rng.standard_normal((N, 3))).astype(np.float32)
    return f, u


def setup(solver):
    solver.set_drive(DRIVE[0], period=DRIVE[1], ramp=DRIVE[2])
    solver.set_force(FORCE)
    solver.set_outlet_pressure(P_OUT)


def main():
    lat = lattice()
    N = lat["nbr"].shape[1]
    f0, u0 = initial_state(N)
    out = {"nbr": lat["nbr"], "q": lat["q"], "uin": lat["uin"], "coords": lat["coords"],
           "n_outlets": np.int32(lat["n_outlets"]), "f0": f0, "u0": u0}
    for name, (model, sigma, bz, steps) in CASES.items():
        o = oracle.OracleSolver(lat, NU, model=model, sigma=sigma, bouzidi=bz, threads=1)
        setup(o)
        o.f[o.cur][:] = f0
        o.vels[o.cur][:] = u0
        done = 0
        for s in steps:
            o.run(s - done)
            done = s
            out[f"{name}_f_{s}"] = o.get_f().copy()
            out[f"{name}_rho_{s}"] = o.rho.copy()
            out[f"{name}_u_{s}"] = o.vels[o.cur].copy()
    np.savez_compressed(os.path.join(HERE, "pipe_small.npz"), **out)
    print(f"wrote pipe_small.npz: {N} cells, keys {sorted(out)}")


if __name__ == "__main__":
    main()
