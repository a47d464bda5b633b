"""Host-side DQMC driver over the GPU chain: the reference's DQMC / run! for LocalSweep.

Follows src/flavors/DQMC/DQMC.jl:
  DQMC(model; kwargs...)   :32-64    parameters, seeded field, stack
  init! / initialize_run   :144-198  hopping exponentials, reverse_build_stack + propagate
  sweep_once! / run!       :200-394  sweep, measurements every measure_rate sweeps,
                                     low-acceptance abort (CANCELLED_LOW_ACCEPTANCE)
  DQMCAnalysis             statistics.jl:9-51 (MagnitudeStats of propagation errors etc.)
  AcceptanceStatistics     updates/scheduler.jl:163-199
The stack itself (every matrix of the chain) lives on the GPU behind the C ABI; this
module only sequences it and gathers the physical Green's function for measurements.
"""
from __future__ import annotations

import enum
import math
import random
import time
from dataclasses import dataclass, field as dc_field

import numpy as np

from .parameters import DQMCParameters
from .measurements import MeasurementSet


class ExitCode(enum.IntEnum):
    """helpers.jl:17-22"""
    SUCCESS = 0
    GENERIC_FAILURE = 1
    CANCELLED_TIME_LIMIT = 2
    CANCELLED_LOW_ACCEPTANCE = 3


class MagnitudeStats:
    """Running min/max/mean of log10|x| (statistics.jl:9-36)."""

    def __init__(self):
        self.count = 0
        self.min = math.inf
        self.max = -math.inf
        self.sum = 0.0

    def push(self, x: float):
        v = math.log10(abs(x)) if x != 0 else -math.inf
        self.count += 1
        self.min = min(self.min, v)
        self.max = max(self.max, v)
        self.sum += v

    def mean(self) -> float:
        return self.sum / self.count if self.count else -math.inf

    def __len__(self):
        return self.count


@dataclass
class DQMCAnalysis:
    imaginary_probability: MagnitudeStats = dc_field(default_factory=MagnitudeStats)
    negative_probability: MagnitudeStats = dc_field(default_factory=MagnitudeStats)
    propagation_error: MagnitudeStats = dc_field(default_factory=MagnitudeStats)
    th_runtime: float = 0.0
    me_runtime: float = 0.0


@dataclass
class AcceptanceStatistics:
    accepted: float = 0.0
    total: int = 0

    @property
    def acceptance_rate(self) -> float:
        return self.accepted / self.total if self.total else 0.0


class LocalSweep:
    """One sweep of local updates: 2M slice visits x N proposals (local_updates.jl:70-84)."""
    name = "LocalSweep"


class DQMC:
    def __init__(self, model, *, seed: int = -1, thermalization: int = 100, sweeps: int = 100,
                 measure_rate: int = 10, last_sweep: int = 0, field=None, precision: str = "fp64",
                 device: int = 0, chain_factory=None, silent: bool = False,
                 device_measurements: bool = False, **kwargs):
        self.parameters = DQMCParameters.create(thermalization=thermalization, sweeps=sweeps,
                                                measure_rate=measure_rate, silent=silent, **kwargs)
        self.model = model
        self.seed = seed if seed != -1 else random.getrandbits(62)
        self.field = (field or model.choose_field())(self.parameters, model)
        self.field.rand(self.seed)
        self.last_sweep = last_sweep
        self.precision = precision
        self.device = device
        self.scheduler = [LocalSweep()]
        self.acceptance = AcceptanceStatistics()
        self.analysis = DQMCAnalysis()
        self.measurements = {}
        self.thermalization_measurements = {}
        self._chain_factory = chain_factory
        # equal-time observables computed on the device (G stays in HBM); the mean Green's
        # function is then not recorded
        self.device_measurements = device_measurements
        self.stack = None

    # ---- stack (drop-in for DQMCStack) ------------------------------------------------
    def _make_chain(self):
        p = self.parameters
        if self._chain_factory is not None:
            return self._chain_factory(len(self.model.l), self.field.flavors, p.slices, p.safe_mult,
                                       p.delta_tau, self.field.alpha)
        from ._lib import HIPChain
        return HIPChain(len(self.model.l), self.field.flavors, p.slices, p.safe_mult, p.delta_tau,
                        self.field.alpha, check_propagation_error=p.check_propagation_error,
                        precision=self.precision, device=self.device)

    def init(self):
        """init! (DQMC.jl:144-150): hopping exponentials and the stack buffers."""
        if self.stack is None:
            self.stack = self._make_chain()
        self.stack.set_hopping(self.model.hopping_matrix())
        self.stack.set_field(self.field.conf)
        self.stack.set_seed((self.seed * 0x9E3779B97F4A7C15 + self.last_sweep) % (1 << 64))

    @property
    def current_slice(self) -> int:
        return self.stack.state()[0]

    def greens(self, physical: bool = True) -> np.ndarray:
        """greens(mc) (greens.jl:87-125): (flavors, N, N)."""
        return self.stack.greens(physical=physical)

    def calculate_greens(self, slice_: int) -> np.ndarray:
        return self.stack.calculate_greens(slice_)

    # ---- run! ---------------------------------------------------------------------------
    def initialize_run(self):
        self.init()
        self.stack.build_stack()
        T = self.model.hopping_matrix()
        self.measurements.setdefault("equal_time", MeasurementSet(T, self.model.U))

    def run(self, *, verbose: bool = True, sweeps: int | None = None,
            thermalization: int | None = None, min_update_rate: float = 0.001) -> ExitCode:
        p = self.parameters
        if sweeps is not None or thermalization is not None:
            self.parameters = p = p.with_sweeps(p.thermalization if thermalization is None else thermalization,
                                                p.sweeps if sweeps is None else sweeps)
        self.initialize_run()
        total = p.thermalization + p.sweeps
        n, M = len(self.model.l), p.slices
        t0 = time.perf_counter()
        max_acc = 0.0
        while self.last_sweep < total:
            acc = self.stack.sweep(1)
            self.last_sweep += 1
            rate = acc / (2 * n * M)
            self.acceptance.accepted += rate
            self.acceptance.total += 1
            max_acc = max(max_acc, rate)
            if self.last_sweep > p.thermalization and self.last_sweep % p.measure_rate == 0:
                if self.device_measurements and hasattr(self.stack, "measure_equal_time"):
                    self.measurements["equal_time"].push_values(self.stack.measure_equal_time(self.model.U))
                else:
                    self.measurements["equal_time"].push(self.greens(physical=True))
            if self.last_sweep == p.thermalization:
                self.analysis.th_runtime += time.perf_counter() - t0
                t0 = time.perf_counter()
            if self.last_sweep >= 1 / min_update_rate and max_acc < min_update_rate:
                if verbose:
                    print(f"Maximum acceptance rate {max_acc} below {min_update_rate}: aborting")
                return ExitCode.CANCELLED_LOW_ACCEPTANCE
            if verbose and self.last_sweep % p.print_rate == 0:
                print(f"[{self.last_sweep}/{total}] acceptance {rate:.3f}")
        self.analysis.me_runtime += time.perf_counter() - t0
        self._collect_errors()
        return ExitCode.SUCCESS

    def _collect_errors(self):
        st = self.stack.counters()
        for x in getattr(self.stack, "stabilization_log", lambda: [])():
            if x > 1e-7:
                self.analysis.propagation_error.push(x)
        if st.get("negative_prob", 0):
            self.analysis.negative_probability.count += st["negative_prob"]


def run(mc: DQMC, **kwargs) -> ExitCode:
    """run!(mc) (DQMC.jl:252-394)."""
    return mc.run(**kwargs)
