"""ctypes binding of libdqmc_hip.so (the C ABI declared in include/dqmc_hip.h).

The library is the only compute path: if it is not built, or no GPU is visible, the
calls raise instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdqmc_hip.so")

PRECISION = {"fp64": 0, "mixed": 1}


class DQMCError(RuntimeError):
    pass


class Config(C.Structure):
    _fields_ = [("n_sites", C.c_int), ("flavors", C.c_int), ("slices", C.c_int),
                ("safe_mult", C.c_int), ("delta_tau", C.c_double), ("alpha", C.c_double),
                ("precision", C.c_int), ("check_propagation_error", C.c_int),
                ("delay_depth", C.c_int), ("device", C.c_int), ("mixed_tolerance", C.c_double)]


class ServerConfig(C.Structure):
    _fields_ = [("device", C.c_int), ("max_active", C.c_int), ("memory_budget", C.c_int64)]


class Request(C.Structure):
    _fields_ = [("chain", Config), ("hopping", C.c_void_p), ("ld_hopping", C.c_int),
                ("field", C.c_void_p), ("ld_field", C.c_int), ("seed", C.c_uint64),
                ("nsweeps", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("accepted", C.c_int64), ("proposed", C.c_int64), ("uniforms_used", C.c_int64),
                ("negative_probability", C.c_int64), ("ties", C.c_int64),
                ("stream_exhausted", C.c_int64), ("propagation_error_count", C.c_int64),
                ("stabilizations", C.c_int64), ("sweeps", C.c_int64),
                ("delayed_flushes", C.c_int64), ("propagation_error_max", C.c_double),
                ("propagation_error_last", C.c_double), ("device_errors", C.c_int64),
                ("precision_violations", C.c_int64)]


EXPORTED = [
    "dqmc_create", "dqmc_destroy", "dqmc_last_error", "dqmc_footprint_bytes",
    "dqmc_set_hopping", "dqmc_set_hopping_exp", "dqmc_get_hopping_exp",
    "dqmc_set_field", "dqmc_get_field", "dqmc_get_field_packed", "dqmc_set_field_packed",
    "dqmc_set_uniforms", "dqmc_set_seed",
    "dqmc_set_decision_log", "dqmc_get_decision_log", "dqmc_build_stack",
    "dqmc_build_stack_forward", "dqmc_reverse_build_stack", "dqmc_propagate",
    "dqmc_sweep_spatial", "dqmc_sweep", "dqmc_batch_sweep", "dqmc_measure_equal_time", "dqmc_get_greens", "dqmc_set_greens",
    "dqmc_calculate_greens", "dqmc_unequal_time_greens", "dqmc_combined_greens",
    "dqmc_density_susceptibility", "dqmc_global_update", "dqmc_get_state", "dqmc_get_stats",
    "dqmc_set_wait_limit", "dqmc_set_sweep_graph", "dqmc_set_slice_launch",
    "dqmc_get_stabilization_log", "dqmc_synchronize", "dqmc_la_udt", "dqmc_la_rdivp",
    "dqmc_la_calc_greens_udt", "dqmc_la_gemm", "dqmc_set_kernel_timing", "dqmc_get_kernel_times",
    "dqmc_server_create", "dqmc_server_submit", "dqmc_server_poll", "dqmc_server_wait",
    "dqmc_server_result", "dqmc_server_status", "dqmc_server_destroy", "dqmc_server_last_error",
    "dqmc_la_udt_noalloc",
]

KERNEL_KINDS = ["gemm_f64", "gemm_f32", "udt", "rdivp", "local", "local_apply", "udt_panel", "udt_update"]

_lib = None
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="F_CONTIGUOUS")
_i8 = np.ctypeslib.ndpointer(dtype=np.int8, flags="F_CONTIGUOUS")
_u64 = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_i32 = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def load():
    """Load the library (raises DQMCError if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise DQMCError(f"{LIB_PATH} is not built; run montecarlo.jl_amd/build.py "
                        "(the HIP library is the only compute path)")
    L = C.CDLL(LIB_PATH)
    vp, i, d, i64 = C.c_void_p, C.c_int, C.c_double, C.c_int64
    sig = {
        "dqmc_create": [C.POINTER(Config), C.POINTER(vp)],
        "dqmc_destroy": [vp],
        "dqmc_footprint_bytes": [C.POINTER(Config), C.POINTER(i64)],
        "dqmc_set_hopping": [vp, _dp, i],
        "dqmc_set_hopping_exp": [vp, _dp, _dp, _dp, _dp, i],
        "dqmc_get_hopping_exp": [vp, _dp, _dp, _dp, _dp, i],
        "dqmc_set_field": [vp, _i8, i],
        "dqmc_get_field": [vp, _i8, i],
        "dqmc_get_field_packed": [vp, _u64, i64],
        "dqmc_set_field_packed": [vp, _u64, i64],
        "dqmc_set_uniforms": [vp, C.POINTER(d), i64],
        "dqmc_set_seed": [vp, C.c_uint64, C.c_uint64],
        "dqmc_set_decision_log": [vp, i64],
        "dqmc_get_decision_log": [vp, vp, vp, i64, C.POINTER(i64)],
        "dqmc_build_stack": [vp], "dqmc_build_stack_forward": [vp],
        "dqmc_reverse_build_stack": [vp], "dqmc_propagate": [vp],
        "dqmc_sweep_spatial": [vp, C.POINTER(i64)],
        "dqmc_sweep": [vp, i, C.POINTER(i64)],
        "dqmc_batch_sweep": [C.POINTER(vp), i, i, C.POINTER(i64)],
        "dqmc_measure_equal_time": [vp, d, vp, i],
        "dqmc_get_greens": [vp, _dp, i, i],
        "dqmc_set_greens": [vp, _dp, i],
        "dqmc_calculate_greens": [vp, i, _dp, i],
        "dqmc_unequal_time_greens": [vp, i, i, _dp, i, i],
        "dqmc_combined_greens": [vp, i, i, i, _dp, i, i],
        "dqmc_density_susceptibility": [vp, i, i, _i32, i, i, i, i, _dp],
        "dqmc_global_update": [vp, _i8, i, C.POINTER(i), C.POINTER(d)],
        "dqmc_get_state": [vp, C.POINTER(i), C.POINTER(i), C.POINTER(i)],
        "dqmc_get_stats": [vp, C.POINTER(Stats)],
        "dqmc_set_wait_limit": [vp, i64],
        "dqmc_set_sweep_graph": [vp, i],
        "dqmc_set_slice_launch": [vp, i],
        "dqmc_get_stabilization_log": [vp, vp, i64, C.POINTER(i64)],
        "dqmc_synchronize": [vp],
        "dqmc_la_udt": [i, _dp, i, _dp, _dp, _dp, _i32],
        "dqmc_la_rdivp": [i, _dp, _dp, _i32],
        "dqmc_la_calc_greens_udt": [i, _dp, _dp, _dp, _dp, _dp, _dp, _dp],
        "dqmc_la_gemm": [i, i, i, _dp, _dp, _dp, i],
        "dqmc_set_kernel_timing": [vp, i],
        "dqmc_get_kernel_times": [vp, vp, vp, i],
        "dqmc_server_create": [C.POINTER(ServerConfig), C.POINTER(vp)],
        "dqmc_server_submit": [vp, C.POINTER(Request), C.POINTER(i64)],
        "dqmc_server_poll": [vp, i64, C.POINTER(i)],
        "dqmc_server_wait": [vp, i64],
        "dqmc_server_result": [vp, i64, vp, i, vp, i, C.POINTER(i64), C.POINTER(Stats), vp],
        "dqmc_server_status": [vp, C.POINTER(i), C.POINTER(i), C.POINTER(i), C.POINTER(i64)],
        "dqmc_server_destroy": [vp],
        "dqmc_la_udt_noalloc": [i, vp, i, vp, vp, vp, vp, vp, vp],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    L.dqmc_last_error.argtypes = [vp]
    L.dqmc_last_error.restype = C.c_char_p
    L.dqmc_server_last_error.argtypes = [vp]
    L.dqmc_server_last_error.restype = C.c_char_p
    _lib = L
    return L


def _check(rc: int, h=None, what: str = ""):
    if rc != 0:
        msg = load().dqmc_last_error(h)
        raise DQMCError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


def _f(a, dtype=np.float64):
    return np.asfortranarray(a, dtype=dtype)


def footprint_bytes(n: int, flavors: int, slices: int, safe_mult: int, precision: str = "fp64") -> int:
    cfg = Config(n, flavors, slices, safe_mult, 0.1, 0.0, PRECISION[precision], 1, 0, 0)
    out = C.c_int64()
    _check(load().dqmc_footprint_bytes(C.byref(cfg), C.byref(out)), None, "dqmc_footprint_bytes")
    return out.value


def batch_sweep(chains, nsweeps: int = 1) -> np.ndarray:
    """Sweep several distinct chains concurrently (dqmc_batch_sweep); returns the accepted
    flips of each chain."""
    k = len(chains)
    hs = (C.c_void_p * k)(*[c._h for c in chains])
    acc = np.zeros(k, dtype=np.int64)
    rc = load().dqmc_batch_sweep(hs, k, nsweeps, acc.ctypes.data_as(C.POINTER(C.c_int64)))
    if rc != 0:
        msgs = [load().dqmc_last_error(c._h) for c in chains]
        raise DQMCError(f"dqmc_batch_sweep failed ({rc}): "
                        + "; ".join(m.decode() for m in msgs if m))
    return acc


REQUEST_STATES = {0: "queued", 1: "running", 2: "done", 3: "failed"}


class Server:
    """Request loop over independent chains (dqmc_server_*): admission by HBM footprint,
    at most `max_active` chains at once, continuous batching."""

    def __init__(self, memory_budget: int, max_active: int = 8, device: int = 0):
        L = load()
        cfg = ServerConfig(device, max_active, memory_budget)
        h = C.c_void_p()
        rc = L.dqmc_server_create(C.byref(cfg), C.byref(h))
        if rc != 0:
            raise DQMCError(f"dqmc_server_create failed ({rc}): {L.dqmc_server_last_error(None).decode()}")
        self._h = h
        self._shapes = {}

    def _c(self, rc, what):
        if rc != 0:
            raise DQMCError(f"{what} failed ({rc}): {load().dqmc_server_last_error(self._h).decode()}")

    def submit(self, T, conf, flavors: int, safe_mult: int, delta_tau: float, alpha: float,
               seed: int, nsweeps: int, precision: str = "fp64", delay_depth: int = 0) -> int:
        T = _f(T)
        conf = _f(conf, np.int8)
        n, M = conf.shape
        cfg = Config(n, flavors, M, safe_mult, delta_tau, alpha, PRECISION[precision], 1,
                     delay_depth, 0)
        rq = Request(cfg, T.ctypes.data, T.shape[0], conf.ctypes.data, n, seed, nsweeps)
        rid = C.c_int64()
        self._c(load().dqmc_server_submit(self._h, C.byref(rq), C.byref(rid)), "dqmc_server_submit")
        self._shapes[rid.value] = (n, M, flavors)
        return rid.value

    def poll(self, rid: int) -> str:
        st = C.c_int()
        self._c(load().dqmc_server_poll(self._h, rid, C.byref(st)), "dqmc_server_poll")
        return REQUEST_STATES[st.value]

    def wait(self, rid: int):
        self._c(load().dqmc_server_wait(self._h, rid), "dqmc_server_wait")

    def result(self, rid: int) -> dict:
        """Field, G (flavors, n, n), accepted flips, stats and latencies of a finished request."""
        n, M, flv = self._shapes.pop(rid)
        conf = np.zeros((n, M), dtype=np.int8, order="F")
        G = np.zeros((n, flv * n), order="F")
        acc = C.c_int64()
        st = Stats()
        lat = np.zeros(2)
        self._c(load().dqmc_server_result(self._h, rid, conf.ctypes.data, n, G.ctypes.data, n,
                                          C.byref(acc), C.byref(st), lat.ctypes.data),
                "dqmc_server_result")
        return {"field": conf, "greens": np.stack([G[:, b * n:(b + 1) * n] for b in range(flv)]),
                "accepted": acc.value, "stats": {k: getattr(st, k) for k, _ in Stats._fields_},
                "latency_s": float(lat[0]), "queued_s": float(lat[1])}

    def status(self) -> dict:
        q, a, p = C.c_int(), C.c_int(), C.c_int()
        u = C.c_int64()
        self._c(load().dqmc_server_status(self._h, C.byref(q), C.byref(a), C.byref(p), C.byref(u)),
                "dqmc_server_status")
        return {"queued": q.value, "active": a.value, "peak_active": p.value, "used_bytes": u.value}

    def close(self):
        if getattr(self, "_h", None):
            load().dqmc_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HIPChain:
    """One DQMC chain resident on the GPU (C-ABI handle).  Green's functions are
    exchanged as (flavors, n, n) arrays."""

    def __init__(self, n: int, flv: int, slices: int, safe_mult: int, delta_tau: float,
                 alpha: float, check_propagation_error: bool = True, precision: str = "fp64",
                 delay_depth: int = 0, device: int = 0, mixed_tolerance: float = 0.0):
        L = load()
        self.n, self.flv, self.M, self.safe_mult = n, flv, slices, safe_mult
        self.delta_tau, self.alpha, self.precision = delta_tau, alpha, precision
        cfg = Config(n, flv, slices, safe_mult, delta_tau, alpha, PRECISION[precision],
                     int(check_propagation_error), delay_depth, device, mixed_tolerance)
        h = C.c_void_p()
        _check(L.dqmc_create(C.byref(cfg), C.byref(h)), None, "dqmc_create")
        self._h = h
        self._u = None

    def close(self):
        if getattr(self, "_h", None):
            load().dqmc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _c(self, rc, what):
        _check(rc, self._h, what)

    def set_hopping(self, T):
        T = _f(T)
        self._c(load().dqmc_set_hopping(self._h, T, T.shape[0]), "dqmc_set_hopping")

    def set_hopping_exp(self, eTmh, eTph, eT2m, eT2p):
        a = [_f(x) for x in (eTmh, eTph, eT2m, eT2p)]
        self._c(load().dqmc_set_hopping_exp(self._h, *a, self.n), "dqmc_set_hopping_exp")

    def get_hopping_exp(self):
        out = [np.zeros((self.n, self.n), order="F") for _ in range(4)]
        self._c(load().dqmc_get_hopping_exp(self._h, *out, self.n), "dqmc_get_hopping_exp")
        return out

    def set_field(self, conf):
        c = _f(conf, np.int8)
        self._c(load().dqmc_set_field(self._h, c, c.shape[0]), "dqmc_set_field")

    def get_field(self):
        c = np.zeros((self.n, self.M), dtype=np.int8, order="F")
        self._c(load().dqmc_get_field(self._h, c, self.n), "dqmc_get_field")
        return c

    def get_field_packed(self):
        """The field as BitArray chunks, packed on the device (fields.jl:331-334)."""
        c = np.zeros((self.n * self.M + 63) // 64, dtype=np.uint64)
        self._c(load().dqmc_get_field_packed(self._h, c, c.size), "dqmc_get_field_packed")
        return c

    def set_field_packed(self, chunks):
        """Set the field from BitArray chunks, unpacked on the device (decompress!)."""
        c = np.ascontiguousarray(chunks, dtype=np.uint64)
        self._c(load().dqmc_set_field_packed(self._h, c, c.size), "dqmc_set_field_packed")

    def set_uniforms(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        self._u = u
        self._c(load().dqmc_set_uniforms(self._h, u.ctypes.data_as(C.POINTER(C.c_double)), len(u)),
                "dqmc_set_uniforms")

    def set_seed(self, seed: int, position: int = 0):
        self._c(load().dqmc_set_seed(self._h, C.c_uint64(seed), C.c_uint64(position)), "dqmc_set_seed")

    def set_log(self, cap: int):
        self._c(load().dqmc_set_decision_log(self._h, cap), "dqmc_set_decision_log")
        self._logcap = cap

    def get_log(self):
        length = C.c_int64()
        self._c(load().dqmc_get_decision_log(self._h, None, None, 0, C.byref(length)), "log")
        n = length.value
        dec = np.zeros(max(n, 1), dtype=np.uint8)
        p = np.zeros(max(n, 1), dtype=np.float64)
        self._c(load().dqmc_get_decision_log(self._h, dec.ctypes.data, p.ctypes.data, n,
                                             C.byref(length)), "log")
        return dec[:n], p[:n]

    def build_stack(self):
        self._c(load().dqmc_build_stack(self._h), "dqmc_build_stack")

    def build_stack_forward(self):
        self._c(load().dqmc_build_stack_forward(self._h), "dqmc_build_stack_forward")

    def reverse_build_stack(self):
        self._c(load().dqmc_reverse_build_stack(self._h), "dqmc_reverse_build_stack")

    def propagate(self):
        self._c(load().dqmc_propagate(self._h), "dqmc_propagate")

    def sweep_spatial(self) -> int:
        a = C.c_int64()
        self._c(load().dqmc_sweep_spatial(self._h, C.byref(a)), "dqmc_sweep_spatial")
        return a.value

    def sweep(self, nsweeps: int = 1) -> int:
        a = C.c_int64()
        self._c(load().dqmc_sweep(self._h, nsweeps, C.byref(a)), "dqmc_sweep")
        return a.value

    def measure_equal_time(self, U: float) -> dict:
        """Equal-time observables computed on the device (dqmc_measure_equal_time); U is the
        reference's coupling (positive: attractive)."""
        n, f = self.n, self.flv
        out = np.zeros(5 + (f + 1) * n)
        self._c(load().dqmc_measure_equal_time(self._h, float(U), out.ctypes.data, out.size),
                "dqmc_measure_equal_time")
        return {"density": float(out[0]), "kinetic_energy": float(out[1]),
                "interaction_energy": float(out[2]), "total_energy": float(out[3]),
                "mean_double_occupancy": float(out[4]),
                "occupation": out[5:5 + f * n].copy(), "double_occupancy": out[5 + f * n:].copy()}

    def _blocks(self, G):
        return np.stack([G[:, b * self.n:(b + 1) * self.n] for b in range(self.flv)])

    def greens(self, physical: bool = False) -> np.ndarray:
        G = np.zeros((self.n, self.n * self.flv), order="F")
        self._c(load().dqmc_get_greens(self._h, G, self.n, int(physical)), "dqmc_get_greens")
        return self._blocks(G)

    def set_greens(self, G):
        flat = np.asfortranarray(np.concatenate(list(G), axis=1))
        self._c(load().dqmc_set_greens(self._h, flat, self.n), "dqmc_set_greens")

    def calculate_greens(self, slice_: int) -> np.ndarray:
        G = np.zeros((self.n, self.n * self.flv), order="F")
        self._c(load().dqmc_calculate_greens(self._h, slice_, G, self.n), "dqmc_calculate_greens")
        return self._blocks(G)

    def unequal_time_greens(self, k: int, l: int, physical: bool = True) -> np.ndarray:
        """greens(mc, k, l) (physical) or calculate_greens(mc, k, l): G(k <- l), (flv, n, n)."""
        G = np.zeros((self.n, self.n * self.flv), order="F")
        self._c(load().dqmc_unequal_time_greens(self._h, k, l, G, self.n, int(physical)),
                "dqmc_unequal_time_greens")
        return self._blocks(G)

    def combined_greens(self, start: int = 0, stop: int | None = None,
                        recalculate: int | None = None):
        """CombinedGreensIterator(mc; start, stop, recalculate) iterated to its end: a list of
        (G(0, l), G(l, 0), G(l, l)) for l = start..stop, each (flv, n, n), physical forms."""
        stop = self.M if stop is None else stop
        rec = 2 * self.safe_mult if recalculate is None else recalculate
        cnt = max(0, stop - start + 1)
        n, f = self.n, self.flv
        out = np.zeros((n, f * n, 3 * cnt), order="F")
        self._c(load().dqmc_combined_greens(self._h, start, stop, rec, out, n, 0),
                "dqmc_combined_greens")
        return [tuple(self._blocks(np.asfortranarray(out[:, :, 3 * i + j])) for j in range(3))
                for i in range(cnt)]

    def density_susceptibility(self, lattice, observable: str = "charge",
                               recalculate: int | None = None, reduced: bool = False,
                               total_flavors: int = 2) -> np.ndarray:
        """charge_density_susceptibility (observable "charge") or
        spin_density_susceptibility (observable "x", "y" or "z") for the current field:
        (Ndir, B, B) with Ndir = Bravais cells of `lattice`, B = its basis sites."""
        code = {"charge": 0, "x": 1, "y": 2, "z": 3}[observable]
        N, B = lattice.n_bravais, lattice.n_basis
        d2st = np.ascontiguousarray(lattice.bravais_dir2srctrg(), dtype=np.int32)
        rec = 2 * self.safe_mult if recalculate is None else recalculate
        out = np.zeros(N * B * B, order="F")
        self._c(load().dqmc_density_susceptibility(self._h, code, rec, d2st, N, B, total_flavors,
                                                   int(reduced), out),
                "dqmc_density_susceptibility")
        return out.reshape((N, B, B), order="F")

    def global_update(self, proposal):
        """global_update with `proposal` as the new conf (dqmc_global_update): (accepted, p)."""
        c = _f(proposal, np.int8)
        acc, p = C.c_int(), C.c_double()
        self._c(load().dqmc_global_update(self._h, c, c.shape[0], C.byref(acc), C.byref(p)),
                "dqmc_global_update")
        return bool(acc.value), p.value

    def state(self):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        self._c(load().dqmc_get_state(self._h, C.byref(a), C.byref(b), C.byref(c)), "state")
        return a.value, b.value, c.value

    def stats(self) -> dict:
        s = Stats()
        self._c(load().dqmc_get_stats(self._h, C.byref(s)), "dqmc_get_stats")
        return {k: getattr(s, k) for k, _ in Stats._fields_}

    def counters(self) -> dict:
        s = self.stats()
        return {"accepted": s["accepted"], "proposed": s["proposed"],
                "uniforms_used": s["uniforms_used"], "negative_prob": s["negative_probability"],
                "ties": s["ties"], "prop_err_count": s["propagation_error_count"],
                "stabilizations": s["stabilizations"], "stream_exhausted": s["stream_exhausted"],
                "prop_err_max": s["propagation_error_max"],
                "prop_err_last": s["propagation_error_last"], "device_errors": s["device_errors"],
                "precision_violations": s["precision_violations"]}

    def set_wait_limit(self, polls: int):
        """Bound of the window kernel's internal wait (dqmc_set_wait_limit)."""
        self._c(load().dqmc_set_wait_limit(self._h, polls), "dqmc_set_wait_limit")

    def set_sweep_graph(self, enable: bool):
        """Replay sweeps from a captured HIP graph (default) or launch kernel by kernel."""
        self._c(load().dqmc_set_sweep_graph(self._h, int(enable)), "dqmc_set_sweep_graph")

    def set_slice_launch(self, enable: bool):
        """Run each slice's windows in one launch (default) or launch once per window; the
        results are bit-identical."""
        self._c(load().dqmc_set_slice_launch(self._h, int(enable)), "dqmc_set_slice_launch")

    def stabilization_log(self) -> np.ndarray:
        length = C.c_int64()
        L = load()
        buf = np.zeros(1 << 16, dtype=np.float64)
        self._c(L.dqmc_get_stabilization_log(self._h, buf.ctypes.data, len(buf), C.byref(length)),
                "dqmc_get_stabilization_log")
        return buf[:min(length.value, len(buf))].copy()

    def synchronize(self):
        self._c(load().dqmc_synchronize(self._h), "dqmc_synchronize")

    def set_kernel_timing(self, enable: bool):
        self._c(load().dqmc_set_kernel_timing(self._h, int(enable)), "dqmc_set_kernel_timing")

    def kernel_times(self) -> dict:
        """{kind: (total_ms, launches)} accumulated since timing was enabled."""
        k = len(KERNEL_KINDS)
        ms = np.zeros(k)
        cnt = np.zeros(k, dtype=np.int64)
        self._c(load().dqmc_get_kernel_times(self._h, ms.ctypes.data, cnt.ctypes.data, k),
                "dqmc_get_kernel_times")
        return {name: (float(ms[i]), int(cnt[i])) for i, name in enumerate(KERNEL_KINDS)}


# ---- kernel-level entry points -----------------------------------------------------------
def la_udt(X, apply_pivot: bool = True):
    n = X.shape[0]
    U = np.zeros((n, n), order="F")
    D = np.zeros((n, 1), order="F")
    T = np.zeros((n, n), order="F")
    piv = np.zeros(n, dtype=np.int32)
    _check(load().dqmc_la_udt(n, _f(X), int(apply_pivot), U, D, T, piv), None, "dqmc_la_udt")
    return U, D[:, 0].copy(), T, piv


def la_rdivp(A, T, pivot):
    n = A.shape[0]
    out = _f(A).copy(order="F")
    _check(load().dqmc_la_rdivp(n, out, _f(T), np.ascontiguousarray(pivot, dtype=np.int32)),
           None, "dqmc_la_rdivp")
    return out


def la_calc_greens_udt(Ul, Dl, Tl, Ur, Dr, Tr):
    n = Ul.shape[0]
    G = np.zeros((n, n), order="F")
    args = [_f(Ul), _f(np.reshape(Dl, (n, 1))), _f(Tl), _f(Ur), _f(np.reshape(Dr, (n, 1))), _f(Tr)]
    _check(load().dqmc_la_calc_greens_udt(n, *args, G), None, "dqmc_la_calc_greens_udt")
    return G


def la_gemm(A, B, ta=False, tb=False, precision: str = "fp64"):
    n = A.shape[0]
    Cm = np.zeros((n, n), order="F")
    _check(load().dqmc_la_gemm(n, int(ta), int(tb), _f(A), _f(B), Cm, PRECISION[precision]),
           None, "dqmc_la_gemm")
    return Cm
