"""cs_flux_bc_static, the flow entry point that allocates nothing, against cs_flux_bc through the C ABI: bitwise
equal flows and opening pressures on two calls in a row, the second with fewer cells and openings (the static
scratch keeps the first call's chunk sums)."""
import ctypes

import numpy as np
import pytest

from cerebrosim_amd import _native

pytestmark = pytest.mark.gpu


def _case(rng, n_cells, m, n_open, torch):
    dev = "cuda"
    vel = torch.from_numpy(rng.standard_normal((n_cells, 3)).astype(np.float32)).to(dev)
    idx = torch.from_numpy(rng.integers(0, n_cells, m).astype(np.int32)).to(dev)
    slot = torch.from_numpy(rng.integers(0, n_open, m).astype(np.int32)).to(dev)
    w = torch.from_numpy(rng.standard_normal((m, 3)).astype(np.float32)).to(dev)
    R = torch.from_numpy(rng.uniform(0.0, 2.0, n_open).astype(np.float32)).to(dev)
    drive = torch.tensor([0.03], dtype=torch.float32, device=dev)
    p0 = rng.uniform(-0.01, 0.01, n_open).astype(np.float32)
    return vel, idx, slot, w, R, drive, p0


def _args(case, n_open, p_out, q_last, work=None, chunks=0):
    vel, idx, slot, w, R, drive, _ = case
    A = _native.FluxArgs()
    A.idx, A.w, A.slot, A.m = idx.data_ptr(), w.data_ptr(), slot.data_ptr(), idx.numel()
    A.vel, A.vel_elem, A.vel_comp = vel.data_ptr(), 3, 1
    A.R, A.drive, A.p_out, A.q_last = R.data_ptr(), drive.data_ptr(), p_out.data_ptr(), q_last.data_ptr()
    A.n_open, A.pref, A.alpha = n_open, 0.002, 0.25
    A.work = work.data_ptr() if work is not None else None
    A.work_chunks = chunks
    return A


def test_static_scratch_flux_equals_allocating_entry_point():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("needs a GPU")
    rng = np.random.default_rng(5)
    stream = torch.cuda.current_stream()
    h = ctypes.c_void_p(stream.cuda_stream)
    for n_cells, m, n_open in ((400_000, 300_000, 5), (5_000, 1_000, 3)):
        case = _case(rng, n_cells, m, n_open, torch)
        p0 = case[-1]
        chunks = _native.flux_chunks(m)
        work = torch.zeros(n_open * chunks, dtype=torch.float64, device="cuda")
        outs = []
        for static in (False, True):
            p_out = torch.from_numpy(p0.copy()).cuda()
            q_last = torch.full((n_open,), float("nan"), dtype=torch.float32, device="cuda")
            if static:
                _native.call("cs_flux_bc_static", ctypes.byref(_args(case, n_open, p_out, q_last)), h)
            else:
                _native.call("cs_flux_bc", ctypes.byref(_args(case, n_open, p_out, q_last, work, chunks)), h)
            torch.cuda.synchronize()
            outs.append((q_last.cpu().numpy(), p_out.cpu().numpy()))
        (qa, pa), (qb, pb) = outs
        print(f"M {m}, {n_open} openings: flows {qb}", flush=True)
        assert np.array_equal(qa.view(np.uint32), qb.view(np.uint32))
        assert np.array_equal(pa.view(np.uint32), pb.view(np.uint32))
        # the flows are the slab sums (independent FP64 check on the host)
        vel, idx, slot, w = (t.cpu().numpy() for t in case[:4])
        expect = np.array([np.sum((vel[idx[slot == k]].astype(np.float64) * w[slot == k]).sum(1)) for k in range(n_open)])
        assert np.allclose(qb, expect, rtol=1e-5, atol=1e-4)


def test_static_scratch_rejects_too_many_openings():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("needs a GPU")
    rng = np.random.default_rng(6)
    case = _case(rng, 100, 50, 2, torch)
    p_out = torch.zeros(65, dtype=torch.float32, device="cuda")
    q_last = torch.zeros(65, dtype=torch.float32, device="cuda")
    with pytest.raises(_native.NativeError, match="n_open 65"):
        _native.call("cs_flux_bc_static", ctypes.byref(_args(case, 65, p_out, q_last)),
                     ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
