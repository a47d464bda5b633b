"""Hirsch Hubbard-Stratonovich fields, src/flavors/DQMC/fields.jl:327-451.

  DensityHirschField   alpha = acosh(exp(+dtau U / 2)), 1 flavour      (fields.jl:363-395)
  MagneticHirschField  alpha = acosh(exp(-dtau U / 2)), 2 flavours     (fields.jl:412-451)
  conf                 Int8 N x M of +-1, column-major                 (fields.jl:329-330)
  compress             BitArray(conf .== 1)                            (fields.jl:331-334)

Only the real-alpha combinations (field matched to the sign of U) are on the accelerated
path; a mismatched field gives a complex alpha and is rejected here.

The reference seeds `conf` from Julia's global RNG (DQMC.jl:48-50), whose stream is not
reproducible outside Julia; this module draws the initial +-1 field from a SplitMix64
stream keyed by the seed (`field_from_seed`).
"""
from __future__ import annotations

import cmath
import math

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def splitmix64(seed: int, k: np.ndarray) -> np.ndarray:
    """k-th output of a SplitMix64 generator started at `seed` (vectorised)."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (k.astype(np.uint64) + np.uint64(1)) * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniforms(seed: int, start: int, count: int) -> np.ndarray:
    """The uniform stream the library draws for the Metropolis test: u_k in [0, 1)."""
    k = np.arange(start, start + count, dtype=np.uint64)
    return (splitmix64(seed, k) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def field_from_seed(seed: int, N: int, M: int) -> np.ndarray:
    """Seeded +-1 Int8 field, shape (N, M), Fortran order."""
    bits = splitmix64(seed ^ 0x5DEECE66D, np.arange(N * M, dtype=np.uint64)) >> np.uint64(63)
    conf = (2 * bits.astype(np.int8) - 1).astype(np.int8)
    return np.asfortranarray(conf.reshape((M, N)).T)


def bits_to_chunks(bits: np.ndarray) -> np.ndarray:
    """A Bool array in Julia's BitArray storage: the column-major entries in uint64 chunks,
    entry k at bit k % 64 of chunk k // 64, the unused high bits of the last chunk zero."""
    flat = np.asarray(bits, dtype=bool).ravel(order="F")
    nch = (flat.size + 63) // 64
    padded = np.zeros(nch * 64, dtype=np.uint64)
    padded[:flat.size] = flat
    weights = np.left_shift(np.uint64(1), np.arange(64, dtype=np.uint64))
    return (padded.reshape(nch, 64) * weights).sum(axis=1, dtype=np.uint64)


def chunks_to_bits(chunks: np.ndarray, shape) -> np.ndarray:
    """Inverse of bits_to_chunks for an array of the given shape (column-major)."""
    chunks = np.asarray(chunks, dtype=np.uint64)
    count = int(np.prod(shape))
    if chunks.size != (count + 63) // 64:
        raise ValueError("chunk count does not match the shape")
    k = np.arange(count, dtype=np.uint64)
    bits = (chunks[k >> np.uint64(6)] >> (k & np.uint64(63))) & np.uint64(1)
    return bits.astype(bool).reshape(shape, order="F")


def _real_alpha(x: complex, name: str) -> float:
    # maybe_to_float, fields.jl:294
    if abs(x.imag) < 10 * np.finfo(float).eps * max(abs(x.real), np.finfo(float).tiny):
        return float(x.real)
    raise ValueError(f"{name}: complex alpha {x} (field does not match the sign of U); "
                     "only real Hirsch couplings are supported")


class AbstractHirschField:
    flavors = 1
    name = "AbstractHirschField"

    def __init__(self, alpha: float, N: int, M: int):
        self.alpha = float(alpha)
        self.conf = np.ones((N, M), dtype=np.int8, order="F")

    def rand(self, seed: int) -> None:
        self.conf[:, :] = field_from_seed(seed, *self.conf.shape)

    def compress(self) -> np.ndarray:
        return self.conf == 1

    def decompress(self, c: np.ndarray) -> None:
        self.conf[:, :] = (2 * np.asarray(c, dtype=np.int8) - 1)

    def compress_chunks(self) -> np.ndarray:
        """compress(f) as Julia stores the BitArray: uint64 chunks (fields.jl:331)."""
        return bits_to_chunks(self.compress())

    def decompress_chunks(self, chunks: np.ndarray) -> None:
        """decompress!(f, c) from BitArray chunks (fields.jl:334)."""
        self.decompress(chunks_to_bits(chunks, self.conf.shape))

    def energy_boson(self) -> float:
        return 0.0


class DensityHirschField(AbstractHirschField):
    flavors = 1
    name = "DensityHirschField"

    def __init__(self, parameters, model, U: float | None = None):
        U = model.U if U is None else U
        a = cmath.acosh(cmath.exp(0.5 * parameters.delta_tau * complex(U)))
        super().__init__(_real_alpha(a, self.name), len(model.l), parameters.slices)

    def energy_boson(self) -> float:
        return self.alpha * float(self.conf.sum())


class MagneticHirschField(AbstractHirschField):
    flavors = 2
    name = "MagneticHirschField"

    def __init__(self, parameters, model, U: float | None = None):
        U = model.U if U is None else U
        a = cmath.acosh(cmath.exp(-0.5 * parameters.delta_tau * complex(U)))
        super().__init__(_real_alpha(a, self.name), len(model.l), parameters.slices)


def hirsch_alpha(delta_tau: float, U_abs: float) -> float:
    """alpha = acosh(exp(dtau |U| / 2)) for a field matched to the sign of U."""
    return math.acosh(math.exp(0.5 * delta_tau * abs(U_abs)))
