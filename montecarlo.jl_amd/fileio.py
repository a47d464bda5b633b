"""Checkpoints, configuration recorders and resume for the DQMC driver.

Follows the reference's
  save / load / resume!        src/FileIO.jl:10-170 (rename / overwrite with a temporary
                               backup, VERSION, "MC" group, RNG state)
  _save / _load of DQMC        src/flavors/DQMC/FileIO.jl:10-74 (Parameters, Analysis, field,
                               configs, last_sweep, Measurements, Model, Scheduler)
  save_rng / load_rng!         src/FileIO.jl:201-222
  ConfigRecorder / Discarder   src/configurations.jl:24-66, 300-336
  field _save / _load          src/flavors/DQMC/fields.jl:298-318 (field name + conf)

Storage.  JLD2 (an HDF5 container) has no reader or writer in this package's environment, so
a checkpoint is a NumPy .npz archive whose keys are the reference file's paths: "VERSION",
"MC/tag", "MC/Parameters/beta", "MC/field/name", "MC/field/conf", "MC/configs/data",
"RNG/seed", ... .  Every value is a plain array or a string; archives are read back with
allow_pickle=False, so loading a file executes nothing from it.  The field, and every
recorded configuration, is stored in the reference's compressed form: BitArray(conf .== 1)
as Julia lays it out (uint64 chunks, column-major entries, least significant bit first).

RNG state.  The reference stores Julia's global RNG.  The library draws the Metropolis
uniforms from a counter stream, u_k = SplitMix64(seed, k); its whole state is (seed, k), so
"RNG/seed" and "RNG/position" restore it exactly.
"""
from __future__ import annotations

import os
import random
import string

import numpy as np

from .fields import bits_to_chunks, chunks_to_bits

FILE_VERSION = 2      # FileIO.jl:36 writes VERSION 2
DQMC_VERSION = 3      # flavors/DQMC/FileIO.jl:11


class ConfigRecorder:
    """Accumulates compressed field configurations every `rate` sweeps of the measurement
    stage (configurations.jl:24-36); `compress` is BitArray(conf .== 1)."""

    def __init__(self, rate: int = 10):
        self.rate = int(rate)
        self.configs: list[np.ndarray] = []   # uint64 BitArray chunks per configuration

    def push(self, field, sweep: int, chunks: np.ndarray | None = None):
        if sweep % self.rate == 0:
            self.configs.append(field.compress_chunks() if chunks is None
                                else np.asarray(chunks, dtype=np.uint64).copy())

    def __len__(self):
        return len(self.configs)

    def __getitem__(self, i):
        return self.configs[i]

    def __iter__(self):
        return iter(self.configs)


class Discarder:
    """Records nothing (configurations.jl:300-336)."""
    rate = 0

    def push(self, field, sweep: int, chunks=None):
        pass

    def __len__(self):
        return 0

    def __iter__(self):
        return iter(())


def _unique_filename(filename: str) -> str:
    """existing.npz -> existing_aJ3c.npz (FileIO.jl:62-75)."""
    if not os.path.isfile(filename):
        return filename
    alphabet = string.digits + string.ascii_uppercase + string.ascii_lowercase
    stem, ext = os.path.splitext(filename)
    stem += "_" + random.choice(alphabet)
    while os.path.isfile(stem + ext):
        stem += random.choice(alphabet)
    return stem + ext


def _entries(mc) -> dict:
    p = mc.parameters
    d: dict[str, np.ndarray] = {}

    def put(key, value):
        d[key] = np.asarray(value)

    put("VERSION", FILE_VERSION)
    put("MC/VERSION", DQMC_VERSION)
    put("MC/tag", "DQMC")
    put("MC/last_sweep", mc.last_sweep)
    for k in ("thermalization", "sweeps", "silent", "check_sign_problem", "check_propagation_error",
              "safe_mult", "delta_tau", "beta", "slices", "measure_rate", "print_rate",
              "checkerboard"):
        put(f"MC/Parameters/{k}", getattr(p, k))
    put("MC/Parameters/tag", "DQMCParameters")
    a = mc.analysis
    put("MC/Analysis/tag", "DQMCAnalysis")
    put("MC/Analysis/th_runtime", a.th_runtime)
    put("MC/Analysis/me_runtime", a.me_runtime)
    for name in ("imaginary_probability", "negative_probability", "propagation_error"):
        s = getattr(a, name)
        put(f"MC/Analysis/{name}/count", s.count)
        put(f"MC/Analysis/{name}/min", s.min)
        put(f"MC/Analysis/{name}/max", s.max)
        put(f"MC/Analysis/{name}/sum", s.sum)
    put("MC/field/tag", "Field")
    put("MC/field/name", mc.field.name)
    put("MC/field/conf", mc.field.compress_chunks())
    put("MC/field/shape", np.array(mc.field.conf.shape, dtype=np.int64))
    rec = mc.recorder
    put("MC/configs/tag", type(rec).__name__)
    put("MC/configs/rate", getattr(rec, "rate", 0))
    data = np.stack(list(rec)) if len(rec) else np.zeros((0, 0), dtype=np.uint64)
    put("MC/configs/data", data.astype(np.uint64))
    m = mc.model
    put("MC/Model/tag", "HubbardModel")
    put("MC/Model/t", m.t)
    put("MC/Model/mu", m.mu)
    put("MC/Model/U", m.U)
    put("MC/Model/l/name", m.l.name)
    put("MC/Model/l/Ls", np.array(m.l.Ls, dtype=np.int64))
    put("MC/Scheduler/tag", "SimpleScheduler")
    put("MC/Scheduler/idx", mc.scheduler.idx)
    put("MC/Scheduler/length", len(mc.scheduler.sequence))
    for k, st in enumerate(mc.scheduler.sequence, start=1):
        put(f"MC/Scheduler/sequence/{k}/tag", "AcceptanceStatistics")
        put(f"MC/Scheduler/sequence/{k}/accepted", st.accepted)
        put(f"MC/Scheduler/sequence/{k}/total", st.total)
        put(f"MC/Scheduler/sequence/{k}/update/tag", st.update.name)
    # the numpy generator of the shuffling global moves (PCG64: 128-bit state and increment)
    st = mc.rng.bit_generator.state
    put("MC/rng/state", np.frombuffer(st["state"]["state"].to_bytes(16, "little"), dtype=np.uint8))
    put("MC/rng/inc", np.frombuffer(st["state"]["inc"].to_bytes(16, "little"), dtype=np.uint8))
    put("MC/rng/has_uint32", st["has_uint32"])
    put("MC/rng/uinteger", st["uinteger"])
    for key, ms in mc.measurements.items():
        for name, arr in ms.state().items():
            put(f"MC/Measurements/ME/{key}/{name}", arr)
    put("MC/device_measurements", bool(mc.device_measurements))
    put("MC/precision", mc.precision)
    put("RNG/seed", np.uint64(mc.seed % (1 << 64)))
    put("RNG/stream_seed", np.uint64(mc.stream_seed))
    put("RNG/position", np.uint64(mc.rng_position))
    return d


def save(filename: str, mc, *, overwrite: bool = False, rename: bool = True) -> str:
    """save(filename, mc; overwrite, rename) (FileIO.jl:10-56); returns the file written."""
    if os.path.isfile(filename) and not overwrite and not rename:
        raise FileExistsError(f'Cannot save because "{filename}" already exists. Consider setting '
                              "`rename = true` to adjust the filename or `overwrite = true` to "
                              "overwrite the file.")
    if os.path.isfile(filename) and not overwrite and rename:
        filename = _unique_filename(filename)
    backup = ""
    if os.path.isfile(filename) and overwrite:
        head, tail = os.path.split(filename)
        backup = _unique_filename(os.path.join(head, "." + tail))
        os.replace(filename, backup)
    mc.sync_from_device()
    entries = _entries(mc)
    try:
        with open(filename, "wb") as fh:
            np.savez(fh, **entries)
    except BaseException:
        if backup and os.path.isfile(backup):
            if os.path.isfile(filename):
                os.remove(filename)
            os.replace(backup, filename)
        raise
    if backup and os.path.isfile(backup):
        os.remove(backup)
    return filename


def _read(filename: str) -> dict:
    with np.load(filename, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def load(filename: str, **dqmc_kwargs):
    """load(filename) (FileIO.jl:86-150): the DQMC object of a checkpoint (stack not built)."""
    from . import lattice as lat
    from .dqmc import DQMC, MagnitudeStats
    from .model import HubbardModel

    d = _read(filename)
    if int(d["VERSION"]) not in (1, 2):
        raise ValueError(f"Failed to load {filename} version {int(d['VERSION'])}")
    if int(d["MC/VERSION"]) > DQMC_VERSION:
        raise ValueError(f"Failed to load DQMC version {int(d['MC/VERSION'])}")
    ctors = {"Chain": lat.Chain, "Square": lat.SquareLattice, "Cubic": lat.CubicLattice,
             "Honeycomb": lat.Honeycomb, "Triangular": lat.TriangularLattice}
    lname = str(d["MC/Model/l/name"])
    l = ctors[lname](*[int(x) for x in d["MC/Model/l/Ls"]])
    model = HubbardModel(l=l, U=float(d["MC/Model/U"]), mu=float(d["MC/Model/mu"]),
                         t=float(d["MC/Model/t"]))
    P = lambda k: d[f"MC/Parameters/{k}"]
    from . import updates as U
    seq = []
    for k in range(1, int(d["MC/Scheduler/length"]) + 1):
        name = str(d[f"MC/Scheduler/sequence/{k}/update/tag"])
        seq.append(getattr(U, name)())
    scheduler = U.SimpleScheduler(*seq)
    scheduler.idx = int(d["MC/Scheduler/idx"])
    for k, st in enumerate(scheduler.sequence, start=1):
        st.accepted = float(d[f"MC/Scheduler/sequence/{k}/accepted"])
        st.total = int(d[f"MC/Scheduler/sequence/{k}/total"])
    recorder = ConfigRecorder(int(d["MC/configs/rate"])) if str(d["MC/configs/tag"]) == "ConfigRecorder" \
        else Discarder()
    mc = DQMC(model, seed=int(d["RNG/seed"]), thermalization=int(P("thermalization")),
              sweeps=int(P("sweeps")), measure_rate=int(P("measure_rate")),
              last_sweep=int(d["MC/last_sweep"]), silent=bool(P("silent")),
              check_sign_problem=bool(P("check_sign_problem")),
              check_propagation_error=bool(P("check_propagation_error")),
              safe_mult=int(P("safe_mult")), print_rate=int(P("print_rate")),
              delta_tau=float(P("delta_tau")), beta=float(P("beta")), slices=int(P("slices")),
              precision=str(d["MC/precision"]), recorder=recorder, scheduler=scheduler,
              device_measurements=bool(d["MC/device_measurements"]), **dqmc_kwargs)
    if mc.field.name != str(d["MC/field/name"]):
        raise ValueError(f"field {d['MC/field/name']} does not match the model's {mc.field.name}")
    shape = tuple(int(x) for x in d["MC/field/shape"])
    mc.field.decompress(chunks_to_bits(d["MC/field/conf"], shape))
    if isinstance(recorder, ConfigRecorder):
        recorder.configs = [row.copy() for row in d["MC/configs/data"]]
    a = mc.analysis
    a.th_runtime = float(d["MC/Analysis/th_runtime"])
    a.me_runtime = float(d["MC/Analysis/me_runtime"])
    for name in ("imaginary_probability", "negative_probability", "propagation_error"):
        s = MagnitudeStats()
        s.count = int(d[f"MC/Analysis/{name}/count"])
        s.min = float(d[f"MC/Analysis/{name}/min"])
        s.max = float(d[f"MC/Analysis/{name}/max"])
        s.sum = float(d[f"MC/Analysis/{name}/sum"])
        setattr(a, name, s)
    st = mc.rng.bit_generator.state
    st["state"]["state"] = int.from_bytes(bytes(d["MC/rng/state"]), "little")
    st["state"]["inc"] = int.from_bytes(bytes(d["MC/rng/inc"]), "little")
    st["has_uint32"] = int(d["MC/rng/has_uint32"])
    st["uinteger"] = int(d["MC/rng/uinteger"])
    mc.rng.bit_generator.state = st
    prefix = "MC/Measurements/ME/"
    groups: dict[str, dict] = {}
    for k, v in d.items():
        if k.startswith(prefix):
            key, name = k[len(prefix):].split("/", 1)
            groups.setdefault(key, {})[name] = v
    for key, state in groups.items():
        mc.restore_measurement(key, state)
    mc.stream_seed = int(d["RNG/stream_seed"])
    mc.rng_position = int(d["RNG/position"])
    return mc


def resume(filename: str, **run_kwargs):
    """resume!(filename; kwargs...) (FileIO.jl:155-170): load, re-initialise, continue run!.
    Returns (mc, exit_code)."""
    mc = load(filename)
    state = mc.run(**run_kwargs)
    return mc, state


__all__ = ["ConfigRecorder", "Discarder", "save", "load", "resume", "bits_to_chunks"]
