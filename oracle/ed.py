"""Exact diagonalisation of the Hubbard model on tiny lattices (numpy).

TEST INFRASTRUCTURE ONLY (statistical oracle for the DQMC sweep).

Restates the reference's ED harness test/ED/ED.jl:
  Fock states        ED.jl:6-62   mode index lin = 2 (site - 1) + substate (1 = up, 2 = down),
                                  fermion sign (-1)^(number of occupied modes below lin)
  Hamiltonian        ED.jl:68-113 H = -t sum_{directed bonds, sigma} c+_{to,sigma} c_{from,sigma}
                                      + U_ED sum_i (n_i_up - 1/2)(n_i_dn - 1/2) - mu sum_i n_i
                                  with U_ED = -model.U (model.U > 0 is attractive) and the same
                                  directed bond list as the hopping matrix (bonds(l, Val(true)))
  expectation values ED.jl:480-511 <O> = sum_k exp(-beta E_k) <k|O|k> / Z
  Greens             ED.jl:145-153, 515-536: G[(s1,i1),(s2,i2)] = <c_{i2 s2} c+_{i1 s1}>
  energy             ED.jl:539-552
Dense matrices are fine: 2x2 has 4^4 = 256 states.
"""
from __future__ import annotations

import numpy as np


def _mode(site: int, spin: int) -> int:
    """0-based mode index of (site, spin) with spin 0 = up, 1 = down (ED.jl:44-62)."""
    return 2 * site + spin


def fermion_ops(nmodes: int):
    """Dense c_m (annihilation) matrices in the occupation basis, Jordan-Wigner signs
    counted over lower modes."""
    dim = 1 << nmodes
    ops = []
    states = np.arange(dim)
    for m in range(nmodes):
        c = np.zeros((dim, dim))
        occ = (states >> m) & 1
        lower = states & ((1 << m) - 1)
        signs = np.array([(-1.0) ** bin(x).count("1") for x in lower])
        src = states[occ == 1]
        c[src ^ (1 << m), src] = signs[occ == 1]
        ops.append(c)
    return ops


class HubbardED:
    def __init__(self, model, beta: float):
        N = len(model.l)
        self.N = N
        self.beta = beta
        nm = 2 * N
        self.c = fermion_ops(nm)
        cd = [x.T for x in self.c]
        self.cd = cd
        dim = 1 << nm
        H = np.zeros((dim, dim))
        for frm, to in model.l.bonds(directed=True):
            for s in (0, 1):
                H -= model.t * cd[_mode(to, s)] @ self.c[_mode(frm, s)]
        U_ed = -model.U
        I = np.eye(dim)
        self.n = [[cd[_mode(i, s)] @ self.c[_mode(i, s)] for s in (0, 1)] for i in range(N)]
        for i in range(N):
            nu, nd = self.n[i]
            H += U_ed * (nu - 0.5 * I) @ (nd - 0.5 * I)
            H -= model.mu * (nu + nd)
        self.H = H
        self.E, self.V = np.linalg.eigh(H)
        w = np.exp(-beta * (self.E - self.E.min()))
        self.w = w / w.sum()

    def expect(self, O: np.ndarray) -> float:
        diag = np.einsum("ik,ij,jk->k", self.V, O, self.V)
        return float(np.dot(self.w, diag))

    def energy(self) -> float:
        return float(np.dot(self.w, self.E))

    def greens(self) -> np.ndarray:
        """2N x 2N, spin-up block first: G[(s1,i1),(s2,i2)] = <c_{i2 s2} c+_{i1 s1}>."""
        N = self.N
        G = np.zeros((2 * N, 2 * N))
        for s1 in (0, 1):
            for s2 in (0, 1):
                for i1 in range(N):
                    for i2 in range(N):
                        O = self.c[_mode(i2, s2)] @ self.cd[_mode(i1, s1)]
                        G[s1 * N + i1, s2 * N + i2] = self.expect(O)
        return G

    def occupation(self) -> np.ndarray:
        """<n_{i,s}>, spin-up sites first."""
        return np.array([self.expect(self.n[i][s]) for s in (0, 1) for i in range(self.N)])

    def double_occupancy(self) -> np.ndarray:
        return np.array([self.expect(self.n[i][0] @ self.n[i][1]) for i in range(self.N)])
