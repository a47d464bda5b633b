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
  integrated <A(tau) B(0)>  ED.jl:659-718 (trapezoid over tau = 0, dtau, .., beta)
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

    def integrated(self, O1: np.ndarray, O2: np.ndarray, step: float) -> float:
        """expectation_value_integrated (ED.jl:659-718): sum over tau = 0, step, .., beta of
        trapezoid weights (step / 2 at the ends) times <O1(tau) O2(0)>, with
        <O1(tau) O2> = sum_nm exp(-(beta - tau) E_n - tau E_m) (O1)_nm (O2)_mn / Z."""
        V, E = self.V, self.E - self.E.min()
        X = (V.T @ O1 @ V) * (V.T @ O2 @ V).T
        Z = np.exp(-self.beta * E).sum()
        M = int(round(self.beta / step))
        out = 0.0
        for i in range(M + 1):
            tau = step * i
            f = np.exp(-(self.beta - tau) * E)[:, None] * np.exp(-tau * E)[None, :]
            out += (0.5 * step if i in (0, M) else step) * float((f * X).sum()) / Z
        return out

    def number(self, site: int) -> np.ndarray:
        """number_operator(site) (ED.jl:722): n_up + n_dn."""
        return self.n[site][0] + self.n[site][1]

    def spin(self, site: int, direction: str) -> tuple:
        """m_x, m_y, m_z of `site` (ED.jl:209-244) as (real matrix A, factor f) with m = f A:
        m_x = c+_up c_dn + c+_dn c_up, m_y = -i (c+_up c_dn - c+_dn c_up), m_z = n_up - n_dn."""
        up, dn = _mode(site, 0), _mode(site, 1)
        X = self.cd[up] @ self.c[dn]
        if direction == "x":
            return X + X.T, 1.0
        if direction == "y":
            return X - X.T, -1j
        return self.n[site][0] - self.n[site][1], 1.0

    def occupation(self) -> np.ndarray:
        """<n_{i,s}>, spin-up sites first."""
        return np.array([self.expect(self.n[i][s]) for s in (0, 1) for i in range(self.N)])

    def double_occupancy(self) -> np.ndarray:
        return np.array([self.expect(self.n[i][0] @ self.n[i][1]) for i in range(self.N)])
