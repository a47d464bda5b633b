"""ctypes wrapper of the CPU FP64 oracle (dqmc_oracle.c).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg as the parity checker / timed CPU baseline; never by the product package.

`OracleChain` exposes the same method names as the product's chain handle
(montecarlo.jl_amd/_lib.py `HIPChain`) so parity tests drive both identically.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "libdqmc_oracle.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="F_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i8 = np.ctypeslib.ndpointer(dtype=np.int8, flags="F_CONTIGUOUS")


def build(force: bool = False) -> str:
    if force or not os.path.exists(_SO) or \
            os.path.getmtime(_SO) < os.path.getmtime(os.path.join(_HERE, "dqmc_oracle.c")):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "all"], check=True,
                       stdout=subprocess.DEVNULL)
    return _SO


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = C.CDLL(_SO)
        vp = C.c_void_p
        L.orc_create.restype = vp
        L.orc_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int]
        L.orc_destroy.argtypes = [vp]
        L.orc_set_hopping.argtypes = [vp, _dp]
        L.orc_set_hopping_exp.argtypes = [vp, _dp, _dp, _dp, _dp]
        L.orc_get_hopping_exp.argtypes = [vp, _dp, _dp, _dp, _dp]
        L.orc_set_field.argtypes = [vp, _i8]
        L.orc_get_field.argtypes = [vp, _i8]
        L.orc_set_uniforms.argtypes = [vp, C.POINTER(C.c_double), C.c_int64]
        L.orc_set_seed.argtypes = [vp, C.c_uint64]
        L.orc_set_seed_at.argtypes = [vp, C.c_uint64, C.c_uint64]
        L.orc_set_log.argtypes = [vp, C.c_void_p, C.c_void_p, C.c_int64]
        L.orc_build_stack_forward.argtypes = [vp]
        L.orc_reverse_build_stack.argtypes = [vp]
        L.orc_propagate.argtypes = [vp]
        L.orc_sweep_spatial.argtypes = [vp]
        L.orc_sweep_spatial.restype = C.c_int64
        L.orc_sweep.argtypes = [vp, C.c_int]
        L.orc_sweep.restype = C.c_int64
        L.orc_calculate_greens_slice.argtypes = [vp, C.c_int, _dp]
        L.orc_get_greens.argtypes = [vp, _dp, C.c_int]
        L.orc_set_greens.argtypes = [vp, _dp]
        L.orc_get_state.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.orc_get_counters.argtypes = [vp, C.POINTER(C.c_int64)]
        L.orc_prop_err_max.argtypes = [vp]
        L.orc_prop_err_max.restype = C.c_double
        L.orc_prop_err_last.argtypes = [vp]
        L.orc_prop_err_last.restype = C.c_double
        L.orc_get_stack_slot.argtypes = [vp, C.c_int, _dp, _dp, _dp]
        L.orc_num_chunks.argtypes = [vp]
        L.orc_num_chunks.restype = C.c_int
        L.orc_udt.argtypes = [C.c_int, _dp, _dp, _dp, _ip, C.c_int]
        L.orc_rdivp.argtypes = [C.c_int, _dp, _dp, _ip]
        L.orc_calc_greens_udt.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_gemm.argtypes = [C.c_int, C.c_int, _dp, C.c_int, _dp, _dp]
        L.orc_expm_sym.argtypes = [C.c_int, _dp, C.c_double, _dp]
        L.orc_uniform_at.argtypes = [C.c_uint64, C.c_uint64]
        L.orc_uniform_at.restype = C.c_double
        L.orc_generate_chunks.argtypes = [C.c_int, C.c_int, _ip, _ip]
        L.orc_generate_chunks.restype = C.c_int
        L.orc_global_update.argtypes = [vp, _i8, C.POINTER(C.c_double)]
        L.orc_global_update.restype = C.c_int
        _lib = L
    return _lib


def _f(a, dtype=np.float64):
    return np.asfortranarray(a, dtype=dtype)


# ---- low-level linear algebra (unit tests) --------------------------------------------
def udt(X: np.ndarray, apply_pivot: bool = True):
    """udt_AVX_pivot! (UDT.jl:216-309): returns U, D, T(input overwritten), pivot (0-based)."""
    n = X.shape[0]
    T = _f(X).copy(order="F")
    U = np.zeros((n, n), order="F")
    D = np.zeros((n, 1), order="F")
    piv = np.zeros(n, dtype=np.int32)
    lib().orc_udt(n, T, U, D, piv, int(apply_pivot))
    return U, D[:, 0].copy(), T, piv


def rdivp(A: np.ndarray, T: np.ndarray, pivot: np.ndarray) -> np.ndarray:
    n = A.shape[0]
    out = _f(A).copy(order="F")
    lib().orc_rdivp(n, out, _f(T), np.ascontiguousarray(pivot, dtype=np.int32))
    return out


def calc_greens_udt(Ul, Dl, Tl, Ur, Dr, Tr) -> np.ndarray:
    n = Ul.shape[0]
    args = [_f(x).copy(order="F") for x in (Ul, np.reshape(Dl, (n, 1)), Tl, Ur, np.reshape(Dr, (n, 1)), Tr)]
    G = np.zeros((n, n), order="F")
    lib().orc_calc_greens_udt(n, *args, G)
    return G


def gemm(A, B, ta=False, tb=False):
    n = A.shape[0]
    Cm = np.zeros((n, n), order="F")
    lib().orc_gemm(n, int(ta), _f(A), int(tb), _f(B), Cm)
    return Cm


def expm_sym(T, c):
    n = T.shape[0]
    out = np.zeros((n, n), order="F")
    lib().orc_expm_sym(n, _f(T), float(c), out)
    return out


def generate_chunks(length, max_chunk):
    s = np.zeros(length + 1, dtype=np.int32)
    e = np.zeros(length + 1, dtype=np.int32)
    k = lib().orc_generate_chunks(length, max_chunk, s, e)
    return list(zip(s[:k].tolist(), e[:k].tolist()))


class OracleChain:
    """One DQMC chain on the CPU oracle.  G blocks are returned as (flv, n, n)."""

    def __init__(self, n: int, flv: int, slices: int, safe_mult: int, delta_tau: float,
                 alpha: float, check_propagation_error: bool = True):
        self.n, self.flv, self.M, self.safe_mult = n, flv, slices, safe_mult
        self.delta_tau, self.alpha = delta_tau, alpha
        self._h = lib().orc_create(n, flv, slices, safe_mult, delta_tau, alpha,
                                   int(check_propagation_error))
        self._keep = []

    def close(self):
        if self._h:
            lib().orc_destroy(self._h)
            self._h = None

    __del__ = close

    def set_hopping(self, T):
        lib().orc_set_hopping(self._h, _f(T))

    def set_hopping_exp(self, eTmh, eTph, eT2m, eT2p):
        lib().orc_set_hopping_exp(self._h, _f(eTmh), _f(eTph), _f(eT2m), _f(eT2p))

    def get_hopping_exp(self):
        n = self.n
        out = [np.zeros((n, n), order="F") for _ in range(4)]
        lib().orc_get_hopping_exp(self._h, *out)
        return out

    def set_field(self, conf):
        lib().orc_set_field(self._h, _f(conf, np.int8))

    def get_field(self):
        c = np.zeros((self.n, self.M), dtype=np.int8, order="F")
        lib().orc_get_field(self._h, c)
        return c

    def set_uniforms(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        self._keep = [u]
        lib().orc_set_uniforms(self._h, u.ctypes.data_as(C.POINTER(C.c_double)), len(u))

    def set_seed(self, seed: int, position: int = 0):
        lib().orc_set_seed_at(self._h, C.c_uint64(seed), C.c_uint64(position))

    def set_log(self, cap: int):
        self._dec = np.zeros(cap, dtype=np.uint8)
        self._p = np.zeros(cap, dtype=np.float64)
        lib().orc_set_log(self._h, self._dec.ctypes.data, self._p.ctypes.data, cap)

    def get_log(self):
        n = self.counters()["log_len"]
        return self._dec[:n].copy(), self._p[:n].copy()

    def build_stack(self):
        """reverse_build_stack + propagate, as initialize_run (DQMC.jl:178-179)."""
        lib().orc_reverse_build_stack(self._h)
        lib().orc_propagate(self._h)

    def build_stack_forward(self):
        lib().orc_build_stack_forward(self._h)
        lib().orc_propagate(self._h)

    def reverse_build_stack(self):
        lib().orc_reverse_build_stack(self._h)

    def propagate(self):
        lib().orc_propagate(self._h)

    def sweep_spatial(self) -> int:
        return int(lib().orc_sweep_spatial(self._h))

    def sweep(self, nsweeps: int = 1) -> int:
        return int(lib().orc_sweep(self._h, nsweeps))

    def greens(self, physical: bool = False) -> np.ndarray:
        G = np.zeros((self.n, self.n * self.flv), order="F")
        lib().orc_get_greens(self._h, G, int(physical))
        return np.stack([G[:, b * self.n:(b + 1) * self.n] for b in range(self.flv)])

    def set_greens(self, G):
        flat = np.asfortranarray(np.concatenate(list(G), axis=1))
        lib().orc_set_greens(self._h, flat)

    def calculate_greens(self, slice_: int) -> np.ndarray:
        G = np.zeros((self.n, self.n * self.flv), order="F")
        lib().orc_calculate_greens_slice(self._h, slice_, G)
        return np.stack([G[:, b * self.n:(b + 1) * self.n] for b in range(self.flv)])

    def global_update(self, proposal):
        """global_update with `proposal` as the new conf (global_updates.jl:191-207): returns
        (accepted, p)."""
        p = C.c_double()
        r = lib().orc_global_update(self._h, _f(proposal, np.int8), C.byref(p))
        if r < 0:
            raise RuntimeError("global update needs current_slice == 1 and direction == 1")
        return bool(r), p.value

    def state(self):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        lib().orc_get_state(self._h, C.byref(a), C.byref(b), C.byref(c))
        return a.value, b.value, c.value

    def counters(self):
        c = (C.c_int64 * 9)()
        lib().orc_get_counters(self._h, c)
        keys = ["accepted", "proposed", "uniforms_used", "negative_prob", "ties",
                "prop_err_count", "stabilizations", "stream_exhausted", "log_len"]
        d = dict(zip(keys, list(c)))
        d["prop_err_max"] = lib().orc_prop_err_max(self._h)
        d["prop_err_last"] = lib().orc_prop_err_last(self._h)
        return d

    def stack_slot(self, s: int):
        n, f = self.n, self.flv
        U = np.zeros((n, n * f), order="F")
        D = np.zeros((n * f, 1), order="F")
        T = np.zeros((n, n * f), order="F")
        lib().orc_get_stack_slot(self._h, s, U, D, T)
        return U, D[:, 0], T
