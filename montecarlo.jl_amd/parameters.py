"""DQMC parameters (src/flavors/DQMC/parameters.jl:33-139) and chunking (stack.jl:154-158)."""
from __future__ import annotations

import math
import warnings
from dataclasses import dataclass, replace
from typing import List, Tuple


def _julia_round(x: float) -> int:
    """Julia's round(Int, x): round half to even (Python's round does the same)."""
    return int(round(x))


def generate_chunks(length: int, max_chunk_size: int) -> List[Tuple[int, int]]:
    """1-based inclusive (first, last) ranges, reference stack.jl:154-158."""
    n_chunks = -(-length // max_chunk_size)
    step = length / n_chunks
    return [(_julia_round((i - 1) * step) + 1, _julia_round(i * step)) for i in range(1, n_chunks + 1)]


@dataclass
class DQMCParameters:
    thermalization: int = 100
    sweeps: int = 100
    silent: bool = False
    check_sign_problem: bool = True
    check_propagation_error: bool = True
    safe_mult: int = 10
    delta_tau: float = 0.1
    beta: float = 1.0
    slices: int = 10
    measure_rate: int = 10
    print_rate: int = 10
    checkerboard: bool = False

    @classmethod
    def create(cls, *, thermalization: int = 100, sweeps: int = 100, silent: bool = False,
               check_sign_problem: bool = True, check_propagation_error: bool = True,
               safe_mult: int = 10, measure_rate: int = 10, warn_round: bool = True,
               print_rate: int | None = None, checkerboard: bool = False, **kw) -> "DQMCParameters":
        """Keyword constructor mirroring parameters.jl:84-139."""
        if checkerboard:
            raise NotImplementedError("checkerboard decomposition is not on the accelerated path")
        if print_rate is None:
            print_rate = max(10, (thermalization + sweeps) // 100)
        keys = set(kw)
        if keys == {"beta"}:
            kw = {"beta": kw["beta"], "delta_tau": 0.1}
            keys = set(kw)
        if keys == {"delta_tau", "beta", "slices"}:
            delta_tau, beta = float(kw["delta_tau"]), float(kw["beta"])
            slices = _julia_round(beta / delta_tau)
            if slices != kw["slices"]:
                raise ValueError(f"Given slices ({kw['slices']}) does not match calculated slices "
                                 f"beta/delta_tau ~ {slices}")
        elif keys == {"beta", "slices"}:
            beta, slices = float(kw["beta"]), int(kw["slices"])
            delta_tau = beta / slices
        elif keys == {"delta_tau", "slices"}:
            delta_tau, slices = float(kw["delta_tau"]), int(kw["slices"])
            beta = delta_tau * slices
        elif keys == {"delta_tau", "beta"}:
            delta_tau, beta = float(kw["delta_tau"]), float(kw["beta"])
            slices = _julia_round(beta / delta_tau)
            if warn_round and not math.isclose(slices, beta / delta_tau, rel_tol=1.5e-8):
                warnings.warn(f"beta/delta_tau = {beta / delta_tau} not an integer. Rounded to {slices}")
        else:
            raise ValueError(f"Invalid keyword arguments to DQMCParameters {kw}")
        return cls(thermalization, sweeps, silent, check_sign_problem, check_propagation_error,
                   safe_mult, delta_tau, beta, slices, measure_rate, print_rate, checkerboard)

    def with_sweeps(self, thermalization: int, sweeps: int) -> "DQMCParameters":
        return replace(self, thermalization=thermalization, sweeps=sweeps)

    @property
    def chunks(self) -> List[Tuple[int, int]]:
        return generate_chunks(self.slices, self.safe_mult)
