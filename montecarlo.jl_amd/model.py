"""Hubbard model: parameters, Hirsch-field choice and the hopping matrix.

Follows src/models/HubbardModel.jl:
  struct / kw constructor  :23-42   (t, mu, U, lattice; positive U is ATTRACTIVE)
  Attractive / Repulsive   :48-55   (Repulsive negates U)
  choose_lattice           :58-66
  choose_field             :83      (MagneticHirschField iff U < 0)
  hopping_matrix           :112-124 (diag(-mu); T[to, from] += -t for each directed bond)
"""
from __future__ import annotations

import warnings

import numpy as np

from .lattice import Chain, CubicLattice, Lattice, SquareLattice


def choose_lattice(dims: int, L: int) -> Lattice:
    if dims == 1:
        return Chain(L)
    if dims == 2:
        return SquareLattice(L)
    return CubicLattice(L, L, L)


class HubbardModel:
    """Hubbard model on `l` with hopping t, chemical potential mu and interaction U.

    Sign convention of the reference: U > 0 is attractive, U < 0 repulsive.
    """

    def __init__(self, L: int | None = None, dims: int | None = None, *, l: Lattice | None = None,
                 U: float = 1.0, mu: float = 0.0, t: float = 1.0):
        if l is None:
            l = choose_lattice(2 if dims is None else dims, 2 if L is None else L)
        if U < 0.0 and mu != 0.0:
            warnings.warn(f"A repulsive Hubbard model with chemical potential mu = {mu} "
                          "will have a sign problem!")
        self.t = float(t)
        self.mu = float(mu)
        self.U = float(U)
        self.l = l

    @property
    def lattice(self) -> Lattice:
        return self.l

    def __repr__(self) -> str:
        kind = "repulsive" if self.U < 0 else "attractive"
        return (f"{kind} Hubbard model (t={self.t}, mu={self.mu}, U={self.U}, "
                f"{self.l.name} with {len(self.l)} sites)")

    def choose_field(self):
        from .fields import DensityHirschField, MagneticHirschField
        return MagneticHirschField if self.U < 0.0 else DensityHirschField

    def hopping_matrix(self) -> np.ndarray:
        """N x N hopping matrix (Fortran order), reference HubbardModel.jl:112-124."""
        N = len(self.l)
        T = np.zeros((N, N), dtype=np.float64, order="F")
        T[np.arange(N), np.arange(N)] = -self.mu
        for frm, to in self.l.bonds(directed=True):
            T[to, frm] += -self.t
        return T


def AttractiveHubbardModel(*args, **kwargs) -> HubbardModel:
    return HubbardModel(*args, **kwargs)


def RepulsiveHubbardModel(*args, **kwargs) -> HubbardModel:
    kwargs = dict(kwargs)
    kwargs["U"] = -kwargs.get("U", 1.0)
    return HubbardModel(*args, **kwargs)


HubbardModelAttractive = AttractiveHubbardModel
HubbardModelRepulsive = RepulsiveHubbardModel
