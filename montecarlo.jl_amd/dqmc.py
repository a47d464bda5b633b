"""Host-side DQMC driver over the GPU chain: the reference's DQMC / run! for LocalSweep.

Follows src/flavors/DQMC/DQMC.jl:
  DQMC(model; kwargs...)   :32-64    parameters, seeded field, recorder, stack
  init! / initialize_run   :144-198  hopping exponentials, reverse_build_stack + propagate
  sweep_once! / run!       :200-394  sweep, recorder push and measurements, low-acceptance
                                     abort (fail save), time-limit save (safe_before,
                                     grace_period), scheduled saves (safe_every)
  replay!                  :418-508  measurements over recorded configurations
  DQMCAnalysis             statistics.jl:9-51 (MagnitudeStats of propagation errors etc.)
  AcceptanceStatistics     updates/scheduler.jl:163-199
The stack itself (every matrix of the chain) lives on the GPU behind the C ABI; this
module only sequences it and gathers the physical Green's function for measurements.
"""
from __future__ import annotations

import datetime as _dt
import enum
import math
import os
import random
import time
from dataclasses import dataclass, field as dc_field

import numpy as np

from .parameters import DQMCParameters
from .measurements import MeasurementSet
from .updates import AcceptanceStatistics, LocalSweep, SimpleScheduler


class ExitCode(enum.IntEnum):
    """helpers.jl:17-22"""
    GENERIC_FAILURE = -1
    SUCCESS = 0
    CANCELLED_TIME_LIMIT = 1
    CANCELLED_LOW_ACCEPTANCE = 2


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


class DQMC:
    def __init__(self, model, *, seed: int = -1, thermalization: int = 100, sweeps: int = 100,
                 measure_rate: int = 10, last_sweep: int = 0, field=None, precision: str = "fp64",
                 device: int = 0, chain_factory=None, silent: bool = False,
                 device_measurements: bool = False, recorder=None, recording_rate: int | None = None,
                 scheduler: SimpleScheduler | None = None, **kwargs):
        from .fileio import ConfigRecorder
        self.parameters = DQMCParameters.create(thermalization=thermalization, sweeps=sweeps,
                                                measure_rate=measure_rate, silent=silent, **kwargs)
        self.model = model
        self.seed = seed if seed != -1 else random.getrandbits(62)
        self.field = (field or model.choose_field())(self.parameters, model)
        self.field.rand(self.seed)
        # RNG state of the Metropolis uniforms: the library's counter stream (seed, position)
        self.stream_seed = (self.seed * 0x9E3779B97F4A7C15 + 0x632BE59BD9B4E019) % (1 << 64)
        self.rng_position = 0
        self.last_sweep = last_sweep
        self.precision = precision
        self.device = device
        self.scheduler = scheduler if scheduler is not None else SimpleScheduler(LocalSweep())
        # proposals of the shuffling global moves (Julia's RNG is not reproducible outside Julia)
        self.rng = np.random.default_rng(self.seed)
        self.analysis = DQMCAnalysis()
        self.recorder = recorder if recorder is not None else \
            ConfigRecorder(measure_rate if recording_rate is None else recording_rate)
        self.measurements = {}
        self.thermalization_measurements = {}
        self._chain_factory = chain_factory
        # equal-time observables computed on the device (G stays in HBM); the mean Green's
        # function is then not recorded
        self.device_measurements = device_measurements
        self.stack = None
        self._restored = {}

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
        """init! (DQMC.jl:144-150): hopping exponentials (init_hopping_matrices, computed from
        the model's hopping matrix) and the stack buffers; the field and the uniform stream's
        state go to the device."""
        if self.stack is None:
            self.stack = self._make_chain()
        self.stack.set_hopping(self.model.hopping_matrix())
        self.stack.set_field(self.field.conf)
        self.stack.set_seed(self.stream_seed, self.rng_position)

    @property
    def acceptance(self) -> AcceptanceStatistics:
        """Acceptance statistics of the first full-sweep update (the LocalSweep)."""
        return next(s for s in self.scheduler.sequence if s.is_full_sweep)

    def sync_from_device(self):
        """Field and RNG position of the device chain -> this object (before saves)."""
        if self.stack is None:
            return
        self.field.conf[:, :] = self.stack.get_field()
        self.rng_position = int(self.stack.counters()["uniforms_used"])

    def restore_measurement(self, key: str, state: dict):
        self._restored[key] = state

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
        if "equal_time" not in self.measurements:
            ms = MeasurementSet(T, self.model.U)
            if "equal_time" in self._restored:
                ms.restore(self._restored.pop("equal_time"))
            self.measurements["equal_time"] = ms

    def _record(self):
        chunks = self.stack.get_field_packed() if hasattr(self.stack, "get_field_packed") else None
        if chunks is None:
            self.field.conf[:, :] = self.stack.get_field()
        self.recorder.push(self.field, self.last_sweep, chunks)

    def _measure(self):
        for key, m in self.measurements.items():
            if key == "equal_time":
                if self.device_measurements and hasattr(self.stack, "measure_equal_time"):
                    m.push_values(self.stack.measure_equal_time(self.model.U))
                else:
                    m.push(self.greens(physical=True))
            else:
                m.measure(self)

    def __setitem__(self, key: str, measurement):
        """mc[:key] = measurement (DQMC_interface.jl): measured every measure_rate sweeps
        after thermalisation."""
        if key == "equal_time":
            raise KeyError("'equal_time' is the built-in equal-time set")
        if key in self._restored:
            measurement.restore(self._restored.pop(key))
        self.measurements[key] = measurement

    def __getitem__(self, key: str):
        return self.measurements[key]

    def run(self, *, verbose: bool = True, sweeps: int | None = None,
            thermalization: int | None = None, min_update_rate: float = 0.001,
            safe_before=None, safe_every=None, grace_period=300.0,
            resumable_filename: str | None = None, overwrite: bool = False,
            fail_filename: str | None = None) -> ExitCode:
        """run!(mc; kwargs...) (DQMC.jl:252-394).  safe_before is a datetime or a Unix time,
        safe_every and grace_period are timedeltas or seconds."""
        from .fileio import save
        p = self.parameters
        if sweeps is not None or thermalization is not None:
            self.parameters = p = p.with_sweeps(p.thermalization if thermalization is None else thermalization,
                                                p.sweeps if sweeps is None else sweeps)
        secs = lambda x: x.total_seconds() if isinstance(x, _dt.timedelta) else float(x)
        deadline = math.inf if safe_before is None else (
            safe_before.timestamp() if isinstance(safe_before, _dt.datetime) else float(safe_before))
        every = math.inf if safe_every is None else secs(safe_every)
        grace = secs(grace_period)
        stamp = _dt.datetime.fromtimestamp(deadline if math.isfinite(deadline) else
                                           time.time() + 100 * 365.25 * 86400)
        tag = stamp.strftime("%d_%b_%Y-%H_%M")
        resumable_filename = resumable_filename or f"resumable_{tag}.npz"
        fail_filename = fail_filename or f"failed_{tag}.npz"

        self.initialize_run()
        self.scheduler.init(self)
        total = p.thermalization + p.sweeps
        t0 = time.perf_counter()
        last_checkpoint = time.time()
        min_sweeps = round(1 / min_update_rate)
        max_sweep_duration = 0.0
        exit_code = ExitCode.SUCCESS
        while self.last_sweep < total:
            ts = time.perf_counter()
            self.scheduler.update(self)        # updates up to a full sweep; last_sweep += 1
            if self.last_sweep <= p.thermalization:
                if self.last_sweep == p.thermalization:
                    self.analysis.th_runtime += time.perf_counter() - t0
                    t0 = time.perf_counter()
            else:
                self._record()
                if self.last_sweep % p.measure_rate == 0:
                    self._measure()
            max_sweep_duration = max(max_sweep_duration, time.perf_counter() - ts)
            # cancel if updates are not accepted (DQMC.jl:289-302): the largest cumulative
            # acceptance over the updates (max_acceptance, scheduler.jl:291-299), checked once
            # more than round(1 / min_update_rate) sweeps ran
            if self.last_sweep > min_sweeps and self.scheduler.max_acceptance() < min_update_rate:
                if verbose:
                    print("Cancelling Simulation due to low acceptance rate:")
                    print(f"\t{self.last_sweep}: {self.scheduler.max_acceptance()}")
                save(fail_filename, self, overwrite=overwrite, rename=False)
                if overwrite and os.path.isfile(resumable_filename):
                    os.remove(resumable_filename)
                exit_code = ExitCode.CANCELLED_LOW_ACCEPTANCE
                break
            if verbose and self.last_sweep % p.print_rate == 0:
                stats = ", ".join(f"{s.name} {s.acceptance_rate:.3f}" for s in self.scheduler.sequence)
                print(f"[{self.last_sweep}/{total}] acceptance: {stats}")
            # early save before the time limit (DQMC.jl:326-343)
            if deadline - time.time() < grace + 2.0 * max_sweep_duration:
                if verbose:
                    print(f"Early save initiated for sweep #{self.last_sweep}.")
                save(resumable_filename, self, overwrite=overwrite, rename=False)
                exit_code = ExitCode.CANCELLED_TIME_LIMIT
                break
            elif time.time() - last_checkpoint > every:
                if verbose:
                    print("Performing scheduled save.")
                last_checkpoint = time.time()
                save(resumable_filename, self, overwrite=overwrite, rename=False)
        if self.last_sweep <= p.thermalization:
            self.analysis.th_runtime += time.perf_counter() - t0
        else:
            self.analysis.me_runtime += time.perf_counter() - t0
        self.sync_from_device()
        self._collect_errors()
        return exit_code

    def replay(self, configurations=None, *, verbose: bool = False, measure_rate: int = 1) -> ExitCode:
        """replay!(mc, configurations) (DQMC.jl:418-508): for each recorded configuration
        (BitArray chunks), set the field, compute G from scratch at slice 0 and measure."""
        from .fields import chunks_to_bits
        configurations = self.recorder if configurations is None else configurations
        self.init()
        self.stack.build_stack_forward() if hasattr(self.stack, "build_stack_forward") else \
            self.stack.build_stack()
        T = self.model.hopping_matrix()
        self.measurements["equal_time"] = MeasurementSet(T, self.model.U)
        configs = list(configurations)
        shape = self.field.conf.shape
        for i in range(self.last_sweep, len(configs), measure_rate):
            self.field.decompress(chunks_to_bits(configs[i], shape))
            self.stack.set_field(self.field.conf)
            G = self.stack.calculate_greens(0)        # effective G at slice 0, from scratch
            eTph, eTmh = self._half_exponentials()
            Gp = np.stack([eTph @ g @ eTmh for g in G])
            self.measurements["equal_time"].push(Gp)
            self.last_sweep = i + 1
        return ExitCode.SUCCESS

    def _half_exponentials(self):
        eTmh, eTph, _, _ = self.stack.get_hopping_exp()
        return eTph, eTmh

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
