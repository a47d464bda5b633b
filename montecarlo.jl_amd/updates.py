"""Updates and the scheduler of the DQMC driver.

Follows src/flavors/DQMC/updates/:
  LocalSweep                    local_updates.jl:70-84   one sweep of local updates
  global moves (propose_conf!)  global_updates.jl:216-415
      GlobalFlip, GlobalShuffle, SpatialShuffle, TemporalShuffle, Denoise, DenoiseFlip,
      StaggeredDenoise
  global_update                 global_updates.jl:191-207 (on the device: dqmc_global_update)
  AcceptanceStatistics          scheduler.jl:163-199
  SimpleScheduler               scheduler.jl:200-300 (update, max_acceptance)

A global move proposes a new field on the host from the current one (the reference's
conf / temp_conf pair); the library computes its acceptance probability from the D vectors
of the two Green's functions and rebuilds the stack when it is accepted.  Shuffles draw from
a numpy Generator seeded by the simulation (Julia's global RNG stream is not reproducible
outside Julia).
"""
from __future__ import annotations

import numpy as np


class AbstractUpdate:
    name = "AbstractUpdate"
    is_full_sweep = False

    def init(self, mc):
        pass


class LocalSweep(AbstractUpdate):
    """One sweep of local updates: 2M slice visits x N proposals (local_updates.jl:70-84).
    Returns the accepted fraction accepted / (2 N M)."""
    name = "LocalSweep"
    is_full_sweep = True

    def __init__(self, n: int = 1):
        self.n = n

    def update(self, mc) -> float:
        N, M = len(mc.model.l), mc.parameters.slices
        acc = mc.stack.sweep(self.n)
        return acc / (2 * N * M * self.n)


class AbstractGlobalUpdate(AbstractUpdate):
    """propose_conf! then global_update (global_updates.jl:216-221)."""

    def propose(self, mc, conf: np.ndarray) -> np.ndarray:
        raise NotImplementedError

    def update(self, mc) -> float:
        conf = mc.stack.get_field()
        proposal = np.asfortranarray(self.propose(mc, conf), dtype=np.int8)
        accepted, p = mc.stack.global_update(proposal)
        if accepted:
            mc.field.conf[:, :] = proposal
        return 1.0 if accepted else 0.0


class GlobalFlip(AbstractGlobalUpdate):
    """Flips the configuration, +-1 -> -+1 (global_updates.jl:232-242)."""
    name = "GlobalFlip"

    def propose(self, mc, conf):
        return -conf


class GlobalShuffle(AbstractGlobalUpdate):
    """shuffle!(conf) over all entries (global_updates.jl:252-268)."""
    name = "GlobalShuffle"

    def propose(self, mc, conf):
        flat = conf.ravel(order="F").copy()
        mc.rng.shuffle(flat)
        return flat.reshape(conf.shape, order="F")


class SpatialShuffle(AbstractGlobalUpdate):
    """new_conf[i, :] = conf[j, :] for a random permutation of the sites
    (global_updates.jl:272-300)."""
    name = "SpatialShuffle"

    def propose(self, mc, conf):
        idx = mc.rng.permutation(conf.shape[0])
        return conf[idx, :]


class TemporalShuffle(AbstractGlobalUpdate):
    """new_conf[:, k] = conf[:, l] for a random permutation of the slices
    (global_updates.jl:304-332)."""
    name = "TemporalShuffle"

    def propose(self, mc, conf):
        idx = mc.rng.permutation(conf.shape[1])
        return conf[:, idx]


def _neighbour_average(mc, conf):
    """conf[i, l] + 2 sum over the bonds i -> j of conf[j, l] (global_updates.jl:350-357)."""
    avg = conf.astype(np.int64).copy()
    for frm, to in mc.model.l.bonds(directed=True):
        avg[frm, :] += 2 * conf[to, :]
    return avg


class Denoise(AbstractGlobalUpdate):
    """Each site takes the sign of its neighbourhood (global_updates.jl:344-359)."""
    name = "Denoise"

    def propose(self, mc, conf):
        return np.sign(_neighbour_average(mc, conf)).astype(np.int8)


class DenoiseFlip(AbstractGlobalUpdate):
    """Each site takes the opposite sign of its neighbourhood (global_updates.jl:369-384)."""
    name = "DenoiseFlip"

    def propose(self, mc, conf):
        return (-np.sign(_neighbour_average(mc, conf))).astype(np.int8)


class StaggeredDenoise(AbstractGlobalUpdate):
    """Denoise with the factor 1 - 2 (i % 2) of the 1-based site index i
    (global_updates.jl:393-415)."""
    name = "StaggeredDenoise"

    def propose(self, mc, conf):
        i1 = np.arange(1, conf.shape[0] + 1)
        stag = (1 - 2 * (i1 % 2))[:, None]
        return (stag * np.sign(_neighbour_average(mc, conf))).astype(np.int8)


class AcceptanceStatistics:
    """accepted (sum of returned acceptances) / total (calls) (scheduler.jl:163-199)."""

    def __init__(self, update: AbstractUpdate):
        self.update = update
        self.accepted = 0.0
        self.total = 0

    @property
    def name(self) -> str:
        return self.update.name

    @property
    def is_full_sweep(self) -> bool:
        return self.update.is_full_sweep

    @property
    def acceptance_rate(self) -> float:
        return self.accepted / max(1, self.total)

    def __call__(self, mc) -> float:
        a = self.update.update(mc)
        self.total += 1
        self.accepted += a
        return a


class SimpleScheduler:
    """Runs its updates in order, cycling; one call of `update` runs updates until a full
    sweep and counts one sweep (scheduler.jl:200-300).  At least one update must be a full
    sweep."""

    def __init__(self, *updates: AbstractUpdate):
        if not updates or not any(u.is_full_sweep for u in updates):
            raise ValueError("The scheduler needs at least one full sweep (e.g. LocalSweep()).")
        self.sequence = [AcceptanceStatistics(u) for u in updates]
        self.idx = -1

    def init(self, mc):
        for s in self.sequence:
            s.update.init(mc)

    def update(self, mc):
        while True:
            self.idx = (self.idx + 1) % len(self.sequence)
            s = self.sequence[self.idx]
            s(mc)
            if s.is_full_sweep:
                break
        mc.last_sweep += 1

    def max_acceptance(self) -> float:
        return max(s.acceptance_rate for s in self.sequence)


__all__ = ["AbstractUpdate", "LocalSweep", "AbstractGlobalUpdate", "GlobalFlip", "GlobalShuffle",
           "SpatialShuffle", "TemporalShuffle", "Denoise", "DenoiseFlip", "StaggeredDenoise",
           "AcceptanceStatistics", "SimpleScheduler"]
