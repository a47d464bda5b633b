"""Lattices: unit cells and the directed-bond iterator used to build the hopping matrix.

Follows MonteCarlo.jl's lattice code:
  unit cells      src/lattices/constructors.jl:1-70 (Chain, Square, Cubic, Honeycomb, Triangular)
  flat indexing   src/lattices/lattice.jl:240-254  (site = x + Lx (y-1) + Lx Ly (basis-1), 1-based)
  bond shifting   src/lattices/lattice.jl:327-357  (_shift_Bravais, periodic wrap with mod1)
  directed bonds  src/lattices/lattice.jl:367-374  (for idx in 1:prod(Ls), for b in unitcell.bonds)
  Bravais shifts  src/lattices/lattice_cache.jl:71-80, 226-242 (Bravais_dir2srctrg / srctrg2dir)

Indices returned by this module are 0-based; `bonds(directed=True)` yields the same
sequence as the reference's `bonds(l, Val(true))` shifted by one.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Sequence, Tuple

import numpy as np


@dataclass(frozen=True)
class Bond:
    frm: int          # basis index (0-based) within the unit cell, or flat site index
    to: int
    uc_shift: Tuple[int, ...]
    label: int = 1


@dataclass(frozen=True)
class UnitCell:
    name: str
    lattice_vectors: Tuple[Tuple[float, ...], ...]
    sites: Tuple[Tuple[float, ...], ...]
    bonds: Tuple[Bond, ...]


class Lattice:
    """Bravais lattice of `Ls` unit cells with periodic boundaries."""

    def __init__(self, unitcell: UnitCell, Ls: Sequence[int]):
        self.unitcell = unitcell
        self.Ls = tuple(int(x) for x in Ls)
        if len(self.Ls) != len(unitcell.lattice_vectors):
            raise ValueError("dimension mismatch between unit cell and Ls")

    @property
    def name(self) -> str:
        return self.unitcell.name

    def __len__(self) -> int:
        return int(np.prod(self.Ls)) * len(self.unitcell.sites)

    def _shift_bravais(self, flat: int, bond: Bond) -> Tuple[int, int]:
        """Reference `_shift_Bravais` (lattice.jl:339-357), 0-based cell index `flat`."""
        out = 0
        f = 1
        rest = flat
        for d, L in enumerate(self.Ls):
            rest, t = divmod(rest, L)
            t = (t + bond.uc_shift[d]) % L
            out += f * t
            f *= L
        return flat + bond.frm * f, out + bond.to * f

    @property
    def n_bravais(self) -> int:
        return int(np.prod(self.Ls))

    @property
    def n_basis(self) -> int:
        return len(self.unitcell.sites)

    def bravais_dir2srctrg(self) -> np.ndarray:
        """`l[:Bravais_dir2srctrg]` (lattice_cache.jl:226-242) as an (Ndir, N) int32 array,
        0-based: row `dir` is the Bravais shift with (column-major) sub-index `dir`, entry `src`
        the Bravais cell reached from `src` by that shift, periodically wrapped."""
        N = self.n_bravais
        out = np.zeros((N, N), dtype=np.int32)
        subs = np.array(np.unravel_index(np.arange(N), self.Ls, order="F"))   # (D, N)
        Ls = np.array(self.Ls)[:, None]
        for shift in range(N):
            trg = (subs + subs[:, shift:shift + 1]) % Ls
            out[shift] = np.ravel_multi_index(tuple(trg), self.Ls, order="F")
        return out

    def bravais_srctrg2dir(self) -> np.ndarray:
        """`l[:Bravais_srctrg2dir]` (lattice_cache.jl:71-80), 0-based (N, N)."""
        d2st = self.bravais_dir2srctrg()
        N = self.n_bravais
        out = np.full((N, N), -1, dtype=np.int32)
        for d in range(N):
            out[np.arange(N), d2st[d]] = d
        return out

    def bonds(self, directed: bool = True) -> List[Tuple[int, int]]:
        """(from, to) pairs, 0-based.  directed=True matches `bonds(l, Val(true))`."""
        if not directed:
            raise NotImplementedError("only the directed iterator is on the DQMC path")
        ncell = int(np.prod(self.Ls))
        return [self._shift_bravais(idx, b) for idx in range(ncell) for b in self.unitcell.bonds]


def Chain(Lx: int) -> Lattice:
    uc = UnitCell("Chain", ((1.0,),), ((0.0,),),
                  (Bond(0, 0, (1,)), Bond(0, 0, (-1,))))
    return Lattice(uc, (Lx,))


def SquareLattice(Lx: int, Ly: int | None = None) -> Lattice:
    Ly = Lx if Ly is None else Ly
    uc = UnitCell("Square", ((1.0, 0.0), (0.0, 1.0)), ((0.0, 0.0),),
                  (Bond(0, 0, (1, 0)), Bond(0, 0, (0, 1)),
                   Bond(0, 0, (-1, 0)), Bond(0, 0, (0, -1))))
    return Lattice(uc, (Lx, Ly))


def CubicLattice(Lx: int, Ly: int | None = None, Lz: int | None = None) -> Lattice:
    Ly = Lx if Ly is None else Ly
    Lz = Lx if Lz is None else Lz
    uc = UnitCell("Cubic", ((1.0, 0, 0), (0, 1.0, 0), (0, 0, 1.0)), ((0.0, 0.0, 0.0),),
                  (Bond(0, 0, (1, 0, 0)), Bond(0, 0, (0, 1, 0)), Bond(0, 0, (0, 0, 1)),
                   Bond(0, 0, (-1, 0, 0)), Bond(0, 0, (0, -1, 0)), Bond(0, 0, (0, 0, -1))))
    return Lattice(uc, (Lx, Ly, Lz))


def Honeycomb(Lx: int, Ly: int | None = None) -> Lattice:
    Ly = Lx if Ly is None else Ly
    s3 = float(np.sqrt(3.0))
    uc = UnitCell("Honeycomb", ((s3 / 2, -0.5), (s3 / 2, 0.5)), ((0.0, 0.0), (1 / s3, 0.0)),
                  (Bond(0, 1, (0, 0)), Bond(0, 1, (-1, 0)), Bond(0, 1, (0, -1)),
                   Bond(1, 0, (0, 0)), Bond(1, 0, (1, 0)), Bond(1, 0, (0, 1))))
    return Lattice(uc, (Lx, Ly))


def TriangularLattice(Lx: int, Ly: int | None = None) -> Lattice:
    Ly = Lx if Ly is None else Ly
    s3 = float(np.sqrt(3.0))
    uc = UnitCell("Triangular", ((s3 / 2, -0.5), (s3 / 2, 0.5)), ((0.0, 0.0),),
                  (Bond(0, 0, (1, 0)), Bond(0, 0, (0, 1)), Bond(0, 0, (-1, 1)),
                   Bond(0, 0, (-1, 0)), Bond(0, 0, (0, -1)), Bond(0, 0, (1, -1))))
    return Lattice(uc, (Lx, Ly))
