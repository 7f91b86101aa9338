"""ctypes binding of libcerebro_lbm.so (declared in include/cerebro_lbm.h).

The library is built in place by ``__graft_entry__.build()`` (cerebrosim_amd/csrc/Makefile). There is no fallback:
if the library is missing or fails to load, every solver entry point raises.
"""
from __future__ import annotations

import ctypes
import gc
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("CEREBRO_LBM_LIB") or os.path.join(HERE, "libcerebro_lbm.so")

c_int, c_i32, c_i64, c_float, c_void_p = ctypes.c_int, ctypes.c_int32, ctypes.c_int64, ctypes.c_float, ctypes.c_void_p
CS_MODEL_TRT, CS_MODEL_REGULARISED = 0, 1
CS_STORAGE_FP32, CS_STORAGE_FP16 = 0, 1
CS_CASE_SHIFT, CS_MAX_CASES, CS_LIST_SKIP = 22, 1024, 0xFFFFFFFF
CS_REDUCE_BLOCKS = 512
CS_REDUCE_WORK = 2 * CS_REDUCE_BLOCKS


class ClockArgs(ctypes.Structure):
    _fields_ = [("tick", c_void_p), ("coef", c_void_p), ("nh", c_i32), ("period", c_i32), ("ramp", c_i32),
                ("drive", c_void_p)]


class StepArgs(ctypes.Structure):
    _fields_ = [("f_in", c_void_p), ("f_out", c_void_p), ("nbr", c_void_p), ("qw", c_void_p), ("uin", c_void_p),
                ("inl", c_void_p), ("p_out", c_void_p), ("drive", c_void_p), ("om_cell", c_void_p),
                ("vel_old", c_void_p), ("vel", c_void_p), ("rho", c_void_p), ("n_cells", c_i64),
                ("force", c_float * 3), ("magic", c_float), ("sigma", c_float), ("bouzidi", c_i32),
                ("model", c_i32), ("galilean", c_i32)]


class Opening(ctypes.Structure):
    _fields_ = [("kind", c_i32), ("centre", ctypes.c_double * 3), ("normal", ctypes.c_double * 3),
                ("radius", ctypes.c_double)]


class GridIn(ctypes.Structure):
    _fields_ = [("phi", c_void_p), ("shape", c_i32 * 3), ("periodic", c_i32 * 3), ("origin", ctypes.c_double * 3),
                ("dx", ctypes.c_double), ("openings", c_void_p), ("n_openings", c_i32), ("slack", ctypes.c_double),
                ("threads", c_i32)]


class ClockCasesArgs(ctypes.Structure):
    _fields_ = [("tick", c_void_p), ("coef", c_void_p), ("nh", c_void_p), ("period", c_void_p), ("ramp", c_void_p),
                ("drive", c_void_p), ("n_cases", c_i32), ("coef_stride", c_i32)]


class ClockTableArgs(ctypes.Structure):
    _fields_ = [("tick", c_void_p), ("coef", c_void_p), ("nh", c_void_p), ("period", c_void_p), ("ramp", c_void_p),
                ("drive", c_void_p), ("n_cases", c_i32), ("coef_stride", c_i32), ("n_steps", c_i32),
                ("drive_stride", c_i32)]


class FluxCasesArgs(ctypes.Structure):
    _fields_ = [("idx", c_void_p), ("w", c_void_p), ("slot", c_void_p), ("start", c_void_p), ("vel", c_void_p),
                ("vel_elem", c_i64), ("vel_comp", c_i64), ("R", c_void_p), ("drive", c_void_p), ("p_out", c_void_p),
                ("q_last", c_void_p), ("n_open", c_void_p), ("pref", c_void_p), ("alpha", c_void_p),
                ("n_cases", c_i32), ("open_stride", c_i32), ("work", c_void_p), ("work_chunks", c_i32)]


class FluxArgs(ctypes.Structure):
    _fields_ = [("idx", c_void_p), ("w", c_void_p), ("slot", c_void_p), ("m", c_i64), ("vel", c_void_p),
                ("vel_elem", c_i64), ("vel_comp", c_i64), ("R", c_void_p), ("drive", c_void_p), ("p_out", c_void_p),
                ("q_last", c_void_p), ("n_open", c_i32), ("pref", c_float), ("alpha", c_float), ("work", c_void_p),
                ("work_chunks", c_i32)]


def flux_chunks(m):
    """Chunks the flow kernels split m slab cells into (max(1, ceil(m / 262144)) cells per thread of 1024-thread
    workgroups), the minimum work_chunks of cs_flux_bc / cs_flux_cases for that lattice."""
    per = max(1, -(-m // (1024 * 256)))
    return -(-m // (1024 * per))


class BinArgs(ctypes.Structure):
    _fields_ = [("start", c_void_p), ("slots", c_void_p), ("n_bins", c_i64), ("rho", c_void_p), ("rho_elem", c_i64),
                ("vel", c_void_p), ("vel_elem", c_i64), ("vel_comp", c_i64), ("p_scale", ctypes.c_double),
                ("p_offset", ctypes.c_double), ("u_scale", ctypes.c_double), ("p_out", c_void_p), ("u_out", c_void_p)]


class GatherArgs(ctypes.Structure):
    _fields_ = [("idx", c_void_p), ("slot", c_void_p), ("m", c_i64), ("rho", c_void_p), ("vel", c_void_p),
                ("vel_elem", c_i64), ("vel_comp", c_i64), ("out", c_void_p), ("umax", c_void_p),
                ("n_probe", c_i32), ("rho_elem", c_i64)]


class WssArgs(ctypes.Structure):
    _fields_ = [("idx", c_void_p), ("c", c_void_p), ("normal", c_void_p), ("n_vert", c_i64), ("k", c_i32),
                ("vel", c_void_p), ("vel_elem", c_i64), ("vel_comp", c_i64), ("scale", c_float), ("tau", c_void_p),
                ("acc_tau", c_void_p), ("acc_mag", c_void_p)]


class SdfTubesArgs(ctypes.Structure):
    _fields_ = [("pts", c_void_p), ("br_start", c_void_p), ("br_count", c_void_p), ("n_branch", c_i32),
                ("cap_p", c_void_p), ("cap_n", c_void_p), ("cap_r", c_void_p), ("blend", c_float),
                ("origin", c_float * 3), ("dx", c_float), ("n", c_i64 * 3), ("out", c_void_p)]


class LatticeIn(ctypes.Structure):
    _fields_ = [("n_cells", c_i64), ("shape", c_i32 * 3), ("periodic", c_i32 * 3), ("coords", c_void_p),
                ("nbr", c_void_p), ("q", c_void_p), ("uin", c_void_p)]


class BricksSizes(ctypes.Structure):
    _fields_ = [("n_bricks", c_i64), ("n_boundary", c_i64), ("n_interior", c_i64), ("n_slots", c_i64),
                ("n_links", c_i64), ("n_side", c_i64)]


class BricksStepArgs(ctypes.Structure):
    _fields_ = [("f", c_void_p), ("nbr", c_void_p), ("list", c_void_p), ("n_list", c_i64), ("bidx", c_void_p),
                ("lmask", c_void_p), ("smask", c_void_p), ("qlink", c_void_p), ("qstart", c_void_p),
                ("uin", c_void_p), ("side", c_void_p), ("sstart", c_void_p), ("loff", c_void_p), ("ldir", c_void_p),
                ("om_cell", c_void_p), ("om_brick", c_void_p), ("omega", c_float), ("p_out", c_void_p),
                ("drive", c_void_p), ("vel_old", c_void_p), ("vel", c_void_p), ("rho", c_void_p),
                ("rho_old", c_void_p), ("n_slots", c_i64), ("force", c_float * 3), ("magic", c_float),
                ("bouzidi", c_i32), ("model", c_i32), ("galilean", c_i32), ("parity", c_i32), ("boundary", c_i32),
                ("storage", c_i32), ("case_omega", c_void_p), ("case_force", c_void_p), ("p_out_stride", c_i32),
                ("variant", c_i32), ("brec", c_void_p), ("hybrid", c_i32), ("sigma", c_float)]


class BricksConvertArgs(ctypes.Structure):
    _fields_ = [("f", c_void_p), ("nbr", c_void_p), ("bidx", c_void_p), ("lmask", c_void_p), ("smask", c_void_p),
                ("side", c_void_p), ("sstart", c_void_p), ("loff", c_void_p), ("slot_of_cell", c_void_p),
                ("n_cells", c_i64),
                ("f_ref", c_void_p),
                ("parity", c_i32), ("storage", c_i32), ("vel", c_void_p), ("rho", c_void_p), ("n_slots", c_i64)]


# exported symbol -> argument types (restype is int for all but cs_last_error)
SIGNATURES = {
    "cs_last_error": [],
    "cs_version": [],
    "cs_stream_create": [ctypes.POINTER(c_void_p)],
    "cs_stream_destroy": [c_void_p],
    "cs_stream_synchronize": [c_void_p],
    "cs_device_synchronize": [],
    "cs_graph_begin": [c_void_p],
    "cs_graph_end": [c_void_p, ctypes.POINTER(c_void_p)],
    "cs_graph_launch": [c_void_p, c_void_p],
    "cs_graph_destroy": [c_void_p],
    "cs_clock": [ctypes.POINTER(ClockArgs), c_void_p],
    "cs_step_indexed": [ctypes.POINTER(StepArgs), c_void_p],
    "cs_flux_bc": [ctypes.POINTER(FluxArgs), c_void_p],
    "cs_flux_bc_static": [ctypes.POINTER(FluxArgs), c_void_p],
    "cs_clock_cases": [ctypes.POINTER(ClockCasesArgs), c_void_p],
    "cs_clock_table": [ctypes.POINTER(ClockTableArgs), c_void_p],
    "cs_flux_cases": [ctypes.POINTER(FluxCasesArgs), c_void_p],
    "cs_bricks_variant_info": [c_i32, c_void_p, c_void_p],
    "cs_gather": [ctypes.POINTER(GatherArgs), c_void_p],
    "cs_reduce_finite": [c_void_p, c_i64, c_void_p, c_void_p, c_void_p],
    "cs_stream_copy": [c_void_p, c_void_p, c_i64, c_void_p],
    "cs_bin_fields": [ctypes.POINTER(BinArgs), c_void_p],
    "cs_wss": [ctypes.POINTER(WssArgs), c_void_p],
    "cs_sdf_tubes": [ctypes.POINTER(SdfTubesArgs), c_void_p],
    "cs_lattice_build": [ctypes.POINTER(GridIn), ctypes.POINTER(c_void_p)],
    "cs_lattice_sizes": [c_void_p, ctypes.POINTER(c_i64), c_void_p, c_i32],
    "cs_lattice_copy": [c_void_p, ctypes.c_char_p, c_void_p],
    "cs_lattice_view": [c_void_p, ctypes.c_char_p, ctypes.POINTER(c_void_p)],
    "cs_lattice_free": [c_void_p],
    "cs_bricks_build": [ctypes.POINTER(LatticeIn), ctypes.POINTER(c_void_p)],
    "cs_bricks_sizes_of": [c_void_p, ctypes.POINTER(BricksSizes)],
    "cs_bricks_copy": [c_void_p, ctypes.c_char_p, c_void_p],
    "cs_bricks_free": [c_void_p],
    "cs_bricks_step": [ctypes.POINTER(BricksStepArgs), c_void_p],
    "cs_bricks_step_pair": [ctypes.POINTER(BricksStepArgs), ctypes.POINTER(BricksStepArgs), c_void_p],
    "cs_bricks_run_pairs": [ctypes.POINTER(BricksStepArgs), c_i32, c_i64, c_i32, c_i64, c_void_p],
    "cs_bricks_slot_table": [c_void_p],
    "cs_bricks_stage_table": [c_void_p],
    "cs_bricks_record_table": [c_void_p],
    "cs_bricks_get_f": [ctypes.POINTER(BricksConvertArgs), c_void_p],
    "cs_bricks_set_f": [ctypes.POINTER(BricksConvertArgs), c_void_p],
}
CS_E_LAYOUT = -3

_lib = None


class NativeError(RuntimeError):
    pass


def load():
    """Load the library (no device work). Raises ImportError when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} not found: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    # torch first: its libraries need "libamdhip64.so" by that name, ours needs "libamdhip64.so.7". Loaded in this
    # order the process has one HIP runtime (torch's copy, whose soname satisfies our library); loaded the other way
    # round it would get two, and the streams torch creates would be foreign to the runtime our kernels launch on.
    import torch  # noqa: F401
    L = ctypes.CDLL(LIB_PATH)
    for name, args in SIGNATURES.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = ctypes.c_char_p if name == "cs_last_error" else c_int
    _lib = L
    return L


def call(name, *args):
    """Call a C-ABI function and raise NativeError with the library's message on failure."""
    rc = getattr(load(), name)(*args)
    if rc != 0:
        msg = load().cs_last_error().decode(errors="replace")
        raise NativeError(f"{name} returned {rc}: {msg}")
    return rc


class Graph:
    """Executable graph captured from launches on one stream (owns the handle)."""

    def __init__(self, handle):
        self.handle = handle

    def launch(self, stream):
        call("cs_graph_launch", self.handle, stream)

    def __del__(self):
        if self.handle is not None and _lib is not None:
            _lib.cs_graph_destroy(self.handle)
            self.handle = None


class capture:
    """Context manager: capture every launch enqueued on `stream` into a Graph (``cap.graph``)."""

    def __init__(self, stream):
        self.stream = stream
        self.graph = None

    def __enter__(self):
        # no cyclic garbage collection while capturing: finalizers of unreachable objects (graphs, events, device
        # memory of dropped solvers and batches) would make runtime calls that a stream capture does not allow
        self._gc = gc.isenabled()
        gc.disable()
        try:
            call("cs_graph_begin", self.stream)
        except BaseException:
            if self._gc:
                gc.enable()
            raise
        return self

    def __exit__(self, exc_type, exc, tb):
        try:
            h = c_void_p()
            if exc_type is not None:
                load().cs_graph_end(self.stream, ctypes.byref(h))   # end the capture, discard the graph
                if h.value:
                    load().cs_graph_destroy(h)
                return False
            call("cs_graph_end", self.stream, ctypes.byref(h))
            self.graph = Graph(h.value)
            return False
        finally:
            if self._gc:
                gc.enable()
