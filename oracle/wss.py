"""ORACLE -- TEST INFRASTRUCTURE ONLY. numpy restatement of shear() of wss.py:53-57 (the per-frame wall shear stress
from the sparse fit coefficients), the checker of the cs_wss kernel."""
import numpy as np


def shear(U, normals, idx, c, mu=3.5e-3):
    u = U[idx]
    ut = u - (u * normals[:, None]).sum(-1, keepdims=True) * normals[:, None]
    return mu * 1e3 * (c[..., None] * ut).sum(1)
