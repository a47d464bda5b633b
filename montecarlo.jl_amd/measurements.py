"""Equal-time measurements from the physical Green's function (host side).

Mirrors the reference's Wick kernels for the Hubbard model:
  occupation      measurements/constructors/occupation.jl:44-55   <n_i> = 1 - G_ii per flavour
  kinetic energy  measurements/constructors/energy.jl:119-150     sum_ij T_ij (I - G)_ij
                  (x 2 for the spin-degenerate density field)
  interaction     models/HubbardModel.jl:160-182                   -U sum_i (G_up_ii - 1/2)(G_dn_ii - 1/2)
  total energy    measurements/constructors/energy.jl:162-166
Double occupancy <n_up n_dn> = (1 - G_up_ii)(1 - G_dn_ii) for one auxiliary-field
configuration (the two spin sectors are independent at fixed field); it is not a
reference measurement and is checked against exact diagonalisation.

`G` is the physical Green's function as (flavors, N, N): one block for the density
field (both spins identical), two blocks (up, down) for the magnetic field.
"""
from __future__ import annotations

import numpy as np


def spin_blocks(G: np.ndarray):
    return (G[0], G[0]) if G.shape[0] == 1 else (G[0], G[1])


def occupation(G: np.ndarray) -> np.ndarray:
    """Per site and flavour block: 1 - G_ii (flavour-major)."""
    return np.concatenate([1.0 - np.diag(g) for g in G])


def density(G: np.ndarray) -> float:
    """<n> per site, both spins."""
    up, dn = spin_blocks(G)
    return float(np.mean(2.0 - np.diag(up) - np.diag(dn)))


def double_occupancy(G: np.ndarray) -> np.ndarray:
    up, dn = spin_blocks(G)
    return (1.0 - np.diag(up)) * (1.0 - np.diag(dn))


def kinetic_energy(G: np.ndarray, T: np.ndarray) -> float:
    I = np.eye(T.shape[0])
    e = sum(float(np.sum(T * (I - g))) for g in G)
    return 2.0 * e if G.shape[0] == 1 else e


def interaction_energy(G: np.ndarray, U: float) -> float:
    up, dn = spin_blocks(G)
    return float(-U * np.sum((np.diag(up) - 0.5) * (np.diag(dn) - 0.5)))


def total_energy(G: np.ndarray, T: np.ndarray, U: float) -> float:
    return kinetic_energy(G, T) + interaction_energy(G, U)


class MeasurementSet:
    """Running means of the equal-time observables (one sample per call)."""

    def __init__(self, T: np.ndarray, U: float):
        self.T, self.U = T, U
        self.count = 0
        self.sums = {}

    def push(self, G: np.ndarray):
        vals = {
            "greens": G.copy(),
            "occupation": occupation(G),
            "double_occupancy": double_occupancy(G),
            "density": density(G),
            "kinetic_energy": kinetic_energy(G, self.T),
            "interaction_energy": interaction_energy(G, self.U),
        }
        vals["total_energy"] = vals["kinetic_energy"] + vals["interaction_energy"]
        for k, v in vals.items():
            self.sums[k] = self.sums.get(k, 0.0) + v
        self.count += 1

    def push_values(self, vals: dict):
        """One sample of already computed observables (the device path: dqmc_measure_equal_time)."""
        keys = ("occupation", "double_occupancy", "density", "kinetic_energy", "interaction_energy",
                "total_energy")
        for k in keys:
            self.sums[k] = self.sums.get(k, 0.0) + vals[k]
        self.count += 1

    def mean(self, key: str):
        return self.sums[key] / max(self.count, 1)

    def state(self) -> dict:
        """Arrays that restore the running means (checkpoints)."""
        out = {"count": np.asarray(self.count)}
        for k, v in self.sums.items():
            out["sum/" + k] = np.asarray(v)
        return out

    def restore(self, state: dict):
        self.count = int(state["count"])
        self.sums = {k[4:]: (float(v) if np.ndim(v) == 0 else np.array(v))
                     for k, v in state.items() if k.startswith("sum/")}


class GreensAt:
    """GreensAt(k, l) (measurements/greens_iterators.jl:65-77): the time-displaced G(k <- l)."""

    def __init__(self, k: int, l: int | None = None):
        self.k, self.l = k, (k if l is None else l)


class TimeIntegral:
    """TimeIntegral(recalculate) (measurements/greens_iterators.jl:83-104); -1 = 2 safe_mult."""

    def __init__(self, recalculate: int = -1):
        self.recalculate = recalculate


class DeviceMeasurement:
    """A measurement evaluated by the library for the chain's current field, once per
    measure call, averaged like the reference's observables (LogBinner means)."""

    def __init__(self, name: str, evaluate):
        self.name = name
        self._evaluate = evaluate
        self.count = 0
        self.sum = None

    def measure(self, mc):
        v = np.asarray(self._evaluate(mc))
        self.sum = v.copy() if self.sum is None else self.sum + v
        self.count += 1

    def mean(self):
        return self.sum / max(self.count, 1)

    def state(self) -> dict:
        return {"count": np.asarray(self.count), "sum": np.asarray(0.0 if self.sum is None else self.sum)}

    def restore(self, state: dict):
        self.count = int(state["count"])
        self.sum = None if self.count == 0 else np.array(state["sum"])


def _recalc(mc, it):
    return None if it.recalculate == -1 else it.recalculate


def greens_measurement(mc, model=None, greens_iterator=None):
    """greens_measurement(mc, model, GreensAt(k, l)) (constructors/greens.jl): the mean of
    greens(mc, k, l), (flavours, N, N).  Without an iterator: the equal-time G."""
    if greens_iterator is None:
        return DeviceMeasurement("G", lambda m: m.greens(physical=True))
    k, l = greens_iterator.k, greens_iterator.l
    return DeviceMeasurement(f"G({k},{l})", lambda m: m.stack.unequal_time_greens(k, l))


def charge_density_susceptibility(mc, model=None, *, reduced: bool = False,
                                  greens_iterator: TimeIntegral | None = None):
    """charge_density_susceptibility(mc, model) (constructors/charge_density.jl:47) with
    EachSitePairByDistance: (Ndir, B, B) per measurement; reduced = reduced_cdc_kernel."""
    it = greens_iterator or TimeIntegral()
    return DeviceMeasurement("CDS", lambda m: m.stack.density_susceptibility(
        m.model.l, "charge", recalculate=_recalc(m, it), reduced=reduced))


def spin_density_susceptibility(mc, model=None, dir: str = "z", *, reduced: bool = False,
                                greens_iterator: TimeIntegral | None = None):
    """spin_density_susceptibility(mc, model, dir) (constructors/spin_density.jl:54)."""
    if dir not in ("x", "y", "z"):
        raise ValueError("dir must be :x, :y or :z")
    it = greens_iterator or TimeIntegral()
    return DeviceMeasurement(f"SDS{dir}", lambda m: m.stack.density_susceptibility(
        m.model.l, dir, recalculate=_recalc(m, it), reduced=reduced))
