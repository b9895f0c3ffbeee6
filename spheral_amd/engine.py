"""ctypes binding of libspheral_amd.so (the C ABI in include/spx_abi.h).

The engine runs only on the GPU: if the library is missing or no HIP device is
visible, ``lib()`` / ``Run(...)`` raise -- there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

from .abi import SpxConfig

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libspheral_amd.so")

# Every C entry point declared by include/spx_abi.h (checked by the CPU test suite).
EXPORTS = (
    "spx_last_error", "spx_config_default", "spx_device_count", "spx_table_eval", "spx_table_coefficients",
    "spx_run_create", "spx_run_destroy", "spx_run_set_state", "spx_run_update_state", "spx_run_get_field",
    "spx_run_initialize",
    "spx_run_iterate_ideal_h",
    "spx_run_step", "spx_run_advance", "spx_run_time", "spx_run_update_connectivity", "spx_run_neighbor_counts",
    "spx_run_neighbors", "spx_run_sum_density", "spx_run_update_eos", "spx_run_post_state_update",
    "spx_run_evaluate_derivatives", "spx_run_dt", "spx_run_compatible_energy", "spx_run_conservation",
    "spx_run_kernel_times", "spx_run_num_pairs", "spx_run_num_pairs_noalloc", "spx_run_neighbor_builds",
    "spx_batch_create", "spx_batch_destroy", "spx_batch_add_run", "spx_batch_remove_run", "spx_batch_step",
    "spx_batch_run_status", "spx_batch_get_field", "spx_batch_conservation", "spx_batch_usage",
    "spx_batch_kernel_times", "spx_batch_neighbor_builds", "spx_bandwidth_probe",
    "spx_run_set_step_graphs", "spx_run_set_end_of_step_omega", "spx_run_set_group_sizes", "spx_run_step_stats",
    "spx_run_gravity_interactions",
    "spx_batch_set_step_graphs", "spx_batch_set_group_sizes", "spx_batch_step_stats",
    "spx_run_device_bytes", "spx_batch_device_bytes",
)

RUN_ACTIVE = 1
RUN_DONE = 2

_lib = None


class SpxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"spx error {code}: {msg}")
        self.code = code


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run `python -m spheral_amd.build`")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        D = C.POINTER(C.c_double)
        I64 = C.POINTER(C.c_int64)
        I32 = C.POINTER(C.c_int32)
        cfgp = C.POINTER(SpxConfig)
        sig = {
            "spx_last_error": (C.c_char_p, []),
            "spx_config_default": (C.c_int, [C.c_int32, cfgp]),
            "spx_device_count": (C.c_int, [I32]),
            "spx_table_eval": (C.c_int, [cfgp, D, C.c_int64, D, D, D]),
            "spx_table_coefficients": (C.c_int, [cfgp, I64, D, D, D]),
            "spx_run_create": (C.c_int, [cfgp, C.c_int64, C.c_int32, C.POINTER(P)]),
            "spx_run_destroy": (None, [P]),
            "spx_run_set_state": (C.c_int, [P, D, D, D, D, D, D]),
            "spx_run_update_state": (C.c_int, [P, D, D, D, D, D, D]),
            "spx_run_get_field": (C.c_int, [P, C.c_char_p, D, C.c_int64, I64]),
            "spx_run_initialize": (C.c_int, [P]),
            "spx_run_iterate_ideal_h": (C.c_int, [P, C.c_int32, C.c_double, C.POINTER(C.c_int32)]),
            "spx_run_step": (C.c_int, [P, C.c_double, D]),
            "spx_run_advance": (C.c_int, [P, C.c_double, C.c_int64, I64, D]),
            "spx_run_time": (C.c_int, [P, D, I64, D]),
            "spx_run_update_connectivity": (C.c_int, [P]),
            "spx_run_neighbor_counts": (C.c_int, [P, I32]),
            "spx_run_neighbors": (C.c_int, [P, I64, I64, I32, C.c_int64]),
            "spx_run_sum_density": (C.c_int, [P]),
            "spx_run_update_eos": (C.c_int, [P]),
            "spx_run_post_state_update": (C.c_int, [P]),
            "spx_run_evaluate_derivatives": (C.c_int, [P, C.c_double, C.c_double]),
            "spx_run_dt": (C.c_int, [P, D, D]),
            "spx_run_compatible_energy": (C.c_int, [P, C.c_double]),
            "spx_run_conservation": (C.c_int, [P, D]),
            "spx_run_kernel_times": (C.c_int, [P, D, C.c_int32, I32, C.c_char_p, C.c_int64]),
            "spx_run_num_pairs": (C.c_int, [P, I64]),
            "spx_run_num_pairs_noalloc": (C.c_int, [P, I64]),
            "spx_run_neighbor_builds": (C.c_int, [P, I64]),
            "spx_batch_create": (C.c_int, [cfgp, C.c_int64, C.c_int32, C.c_int32, C.POINTER(P)]),
            "spx_batch_destroy": (None, [P]),
            "spx_batch_add_run": (C.c_int, [P, D, D, D, D, D, D, C.c_int64, C.c_double, I32]),
            "spx_batch_remove_run": (C.c_int, [P, C.c_int32]),
            "spx_batch_step": (C.c_int, [P, I32]),
            "spx_batch_run_status": (C.c_int, [P, C.c_int32, I32, D, I64, D, I64]),
            "spx_batch_get_field": (C.c_int, [P, C.c_int32, C.c_char_p, D, C.c_int64, I64]),
            "spx_batch_conservation": (C.c_int, [P, C.c_int32, D]),
            "spx_batch_usage": (C.c_int, [P, I64, I32]),
            "spx_batch_neighbor_builds": (C.c_int, [P, C.c_int32, I64]),
            "spx_run_set_step_graphs": (C.c_int, [P, C.c_int32]),
            "spx_run_set_end_of_step_omega": (C.c_int, [P, C.c_int32]),
            "spx_run_gravity_interactions": (C.c_int, [P, I64]),
            "spx_run_set_group_sizes": (C.c_int, [P, I32]),
            "spx_run_step_stats": (C.c_int, [P, I64]),
            "spx_batch_set_step_graphs": (C.c_int, [P, C.c_int32]),
            "spx_batch_set_group_sizes": (C.c_int, [P, I32]),
            "spx_batch_step_stats": (C.c_int, [P, I64]),
            "spx_run_device_bytes": (C.c_int, [P, I64]),
            "spx_batch_device_bytes": (C.c_int, [P, I64]),
            "spx_batch_kernel_times": (C.c_int, [P, D, C.c_int32, I32, C.c_char_p, C.c_int64]),
            "spx_bandwidth_probe": (C.c_int, [C.c_int32, C.c_int64, C.c_int32, D]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(code):
    if code != 0:
        raise SpxError(code, lib().spx_last_error().decode())


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def bandwidth_probe(nbytes=1 << 30, reps=10, device=0):
    """Achieved HBM copy rate (read + write) of a plain streaming kernel, GB/s."""
    g = C.c_double()
    _check(lib().spx_bandwidth_probe(int(device), int(nbytes), int(reps), C.byref(g)))
    return g.value


def device_count():
    n = C.c_int32(0)
    lib().spx_device_count(C.byref(n))
    return n.value


FIELD_KIND = {
    "position": "v", "velocity": "v", "mass": "s", "mass density": "s", "specific thermal energy": "s",
    "H": "t", "pressure": "s", "sound speed": "s", "grad h corrections": "s", "delta velocity hydro": "v",
    "delta velocity": "v", "delta position": "v", "delta specific thermal energy": "s",
    "delta mass density": "s", "velocity gradient": "t", "delta H": "t", "new H": "t",
    "max viscous pressure": "s", "effective viscous pressure": "s", "mass zeroth moment": "s",
    "mass first moment": "v", "M SPH correction": "t", "mass density gradient": "v",
    "normalization": "s", "new mass density": "s", "gravitational potential": "s",
}



STEP_STATS = ("graph_captures", "graph_replays", "eager_steps", "group_tunings", "group_filter", "group_gradh",
              "group_update", "full_reorders", "in_place_retirements")


def _sizes(sizes):
    if sizes is None:
        return None
    arr = (C.c_int32*3)(*[int(v) for v in sizes])
    return C.cast(arr, C.POINTER(C.c_int32))


def _stats(fn, h):
    out = (C.c_int64*len(STEP_STATS))()
    _check(fn(h, out))
    return dict(zip(STEP_STATS, list(out)))


class Run:
    """One resident simulation (spx_run) on a GPU."""

    def __init__(self, cfg: SpxConfig, n, device=0):
        self.cfg = cfg.copy()
        self.n = int(n)
        self.dim = cfg.dim
        h = C.c_void_p()
        _check(lib().spx_run_create(C.byref(self.cfg), self.n, int(device), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().spx_run_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, x, v, m, rho, eps, H):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, v, m, rho, eps, H)]
        _check(lib().spx_run_set_state(self.h, *[_dp(a) for a in arrs]))

    def update_state(self, x, v, m, rho, eps, H):
        """Replace the run's state in place (spx_run_update_state): derivatives, Omega and the
        connectivity are kept, pressure and sound speed follow from the new density and energy."""
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, v, m, rho, eps, H)]
        _check(lib().spx_run_update_state(self.h, *[_dp(a) for a in arrs]))

    def field(self, name):
        k = FIELD_KIND[name]
        shape = (self.n,) if k == "s" else (self.n, self.dim) if k == "v" else (self.n, self.dim, self.dim)
        out = np.zeros(shape, np.float64)
        cnt = C.c_int64()
        _check(lib().spx_run_get_field(self.h, name.encode(), _dp(out), out.size, C.byref(cnt)))
        return out

    def initialize(self):
        _check(lib().spx_run_initialize(self.h))

    def iterate_ideal_h(self, max_iterations=50, tolerance=1.0e-4):
        """SpheralController.iterateIdealH at start-up; returns the number of passes."""
        k = C.c_int32()
        _check(lib().spx_run_iterate_ideal_h(self.h, int(max_iterations), float(tolerance), C.byref(k)))
        return k.value

    def step(self, max_time=1.0e300):
        dt = C.c_double()
        _check(lib().spx_run_step(self.h, float(max_time), C.byref(dt)))
        return dt.value

    def advance(self, max_time, nsteps):
        k = C.c_int64()
        t = C.c_double()
        _check(lib().spx_run_advance(self.h, float(max_time), int(nsteps), C.byref(k), C.byref(t)))
        return k.value, t.value

    def time(self):
        t, c, l = C.c_double(), C.c_int64(), C.c_double()
        _check(lib().spx_run_time(self.h, C.byref(t), C.byref(c), C.byref(l)))
        return t.value, c.value, l.value

    def update_connectivity(self):
        _check(lib().spx_run_update_connectivity(self.h))

    def neighbor_counts(self):
        out = np.zeros(self.n, np.int32)
        _check(lib().spx_run_neighbor_counts(self.h, out.ctypes.data_as(C.POINTER(C.c_int32))))
        return out

    def neighbors(self):
        counts = self.neighbor_counts()
        total = int(counts.sum())
        off = np.zeros(self.n + 1, np.int64)
        nbr = np.zeros(max(total, 1), np.int64)
        img = np.zeros(max(total, 1), np.int32)
        _check(lib().spx_run_neighbors(self.h, off.ctypes.data_as(C.POINTER(C.c_int64)),
                                       nbr.ctypes.data_as(C.POINTER(C.c_int64)),
                                       img.ctypes.data_as(C.POINTER(C.c_int32)), total))
        return off, nbr[:total], img[:total]

    def sum_density(self):
        _check(lib().spx_run_sum_density(self.h))

    def update_eos(self):
        _check(lib().spx_run_update_eos(self.h))

    def post_state_update(self):
        _check(lib().spx_run_post_state_update(self.h))

    def evaluate_derivatives(self, t=0.0, dt=0.0):
        _check(lib().spx_run_evaluate_derivatives(self.h, float(t), float(dt)))

    def dt(self):
        a, b = C.c_double(), C.c_double()
        _check(lib().spx_run_dt(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def compatible_energy(self, multiplier):
        _check(lib().spx_run_compatible_energy(self.h, float(multiplier)))

    def conservation(self):
        out = np.zeros(7)
        _check(lib().spx_run_conservation(self.h, _dp(out)))
        return out

    def num_pairs(self):
        n = C.c_int64()
        _check(lib().spx_run_num_pairs(self.h, C.byref(n)))
        return n.value

    def num_pairs_noalloc(self):
        """num_pairs() through the entry point that allocates no device memory."""
        n = C.c_int64()
        _check(lib().spx_run_num_pairs_noalloc(self.h, C.byref(n)))
        return n.value

    def device_bytes(self):
        """Device memory the run holds (bytes)."""
        n = C.c_int64()
        _check(lib().spx_run_device_bytes(self.h, C.byref(n)))
        return n.value

    def gravity_interactions(self):
        """(accepted-cell, point-point) interactions of one TreeGravity walk over the current tree."""
        out = (C.c_int64*2)()
        _check(lib().spx_run_gravity_interactions(self.h, out))
        return int(out[0]), int(out[1])

    def set_step_graphs(self, enable):
        """Replay steady CheapSynchronousRK2 steps from captured HIP graphs (default on)."""
        _check(lib().spx_run_set_step_graphs(self.h, 1 if enable else 0))

    def set_end_of_step_omega(self, eager):
        """Compute the end-of-step grad-h Omega at every step (eager) or when it is first read
        (default); the results are the same bits either way."""
        _check(lib().spx_run_set_end_of_step_omega(self.h, 1 if eager else 0))

    def set_group_sizes(self, sizes=None):
        """Fix the neighbour-group sizes (filter, grad-h, update; each 4 or 8), or None: measure."""
        _check(lib().spx_run_set_group_sizes(self.h, _sizes(sizes)))

    def step_stats(self):
        """Graph captures/replays, eager steps, group-size measurements and the current sizes."""
        return _stats(lib().spx_run_step_stats, self.h)

    def neighbor_builds(self):
        """Candidate-list builds so far (start-up, forced rebuilds, and skin exhaustion)."""
        n = C.c_int64()
        _check(lib().spx_run_neighbor_builds(self.h, C.byref(n)))
        return n.value

    def enable_timing(self):
        _check(lib().spx_run_kernel_times(self.h, None, -1, None, None, 0))

    def kernel_times(self):
        ms = np.zeros(32)
        k = C.c_int32()
        buf = C.create_string_buffer(1024)
        _check(lib().spx_run_kernel_times(self.h, _dp(ms), 32, C.byref(k), buf, 1024))
        names = buf.value.decode().split(",") if k.value else []
        return dict(zip(names, ms[:k.value].tolist()))


class Batch:
    """Many independent runs of one problem type stepped together (spx_batch)."""

    def __init__(self, cfg: SpxConfig, capacity, max_runs, device=0):
        self.cfg = cfg.copy()
        self.dim = cfg.dim
        self.sizes = {}
        h = C.c_void_p()
        _check(lib().spx_batch_create(C.byref(self.cfg), int(capacity), int(max_runs), int(device), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().spx_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_run(self, state, t_end):
        arrs = [np.ascontiguousarray(state[k], dtype=np.float64) for k in ("x", "v", "m", "rho", "eps", "H")]
        n = len(arrs[2])
        rid = C.c_int32()
        _check(lib().spx_batch_add_run(self.h, *[_dp(a) for a in arrs], n, float(t_end), C.byref(rid)))
        self.sizes[rid.value] = n
        return rid.value

    def remove_run(self, rid):
        _check(lib().spx_batch_remove_run(self.h, int(rid)))
        self.sizes.pop(rid, None)

    def step(self):
        k = C.c_int32()
        _check(lib().spx_batch_step(self.h, C.byref(k)))
        return k.value

    def status(self, rid):
        st, t, cyc, ldt, n = C.c_int32(), C.c_double(), C.c_int64(), C.c_double(), C.c_int64()
        _check(lib().spx_batch_run_status(self.h, int(rid), C.byref(st), C.byref(t), C.byref(cyc), C.byref(ldt),
                                          C.byref(n)))
        return {"state": st.value, "t": t.value, "cycle": cyc.value, "last_dt": ldt.value, "n": n.value}

    def field(self, rid, name):
        n = self.sizes[rid]
        k = FIELD_KIND[name]
        shape = (n,) if k == "s" else (n, self.dim) if k == "v" else (n, self.dim, self.dim)
        out = np.zeros(shape, np.float64)
        cnt = C.c_int64()
        _check(lib().spx_batch_get_field(self.h, int(rid), name.encode(), _dp(out), out.size, C.byref(cnt)))
        return out

    def conservation(self, rid):
        out = np.zeros(7)
        _check(lib().spx_batch_conservation(self.h, int(rid), _dp(out)))
        return out

    def usage(self):
        p, r = C.c_int64(), C.c_int32()
        _check(lib().spx_batch_usage(self.h, C.byref(p), C.byref(r)))
        return p.value, r.value

    def device_bytes(self):
        """Device memory the batch holds (bytes)."""
        n = C.c_int64()
        _check(lib().spx_batch_device_bytes(self.h, C.byref(n)))
        return n.value

    def set_step_graphs(self, enable):
        _check(lib().spx_batch_set_step_graphs(self.h, 1 if enable else 0))

    def set_group_sizes(self, sizes=None):
        _check(lib().spx_batch_set_group_sizes(self.h, _sizes(sizes)))

    def step_stats(self):
        return _stats(lib().spx_batch_step_stats, self.h)

    def neighbor_builds(self, rid):
        n = C.c_int64()
        _check(lib().spx_batch_neighbor_builds(self.h, rid, C.byref(n)))
        return n.value

    def enable_timing(self):
        _check(lib().spx_batch_kernel_times(self.h, None, -1, None, None, 0))

    def kernel_times(self):
        ms = np.zeros(32)
        k = C.c_int32()
        buf = C.create_string_buffer(1024)
        _check(lib().spx_batch_kernel_times(self.h, _dp(ms), 32, C.byref(k), buf, 1024))
        names = buf.value.decode().split(",") if k.value else []
        return dict(zip(names, ms[:k.value].tolist()))


def make_run(cfg, state, device=0):
    r = Run(cfg, len(state["m"]), device)
    r.set_state(state["x"], state["v"], state["m"], state["rho"], state["eps"], state["H"])
    return r
