"""Where can FP32 arithmetic enter a DQMC chain and keep max|dG| <= 1e-6?

TEST INFRASTRUCTURE (a numerical study on the CPU; never part of the product path).

The chain is thermalised with the FP64 oracle (dqmc_oracle.c), then each placement of FP32
arithmetic is applied to one stabilisation interval of `safe_mult` slices and the propagated
Green's function is compared with the FP64 from-scratch G of the same field at the interval's
end, the quantity the reference checks at every stabilisation (stack.jl:634-654):

  fp64       all FP64 (the reference's own drift)
  upd32      rank-1 updates G -= (e_i - G[:, i]) (gamma G[i, :]) formed in FP32
  wrap32     wraps G <- B G B^-1 in FP32 (FP32 products, FP32 accumulation)
  tail j     G held and updated in FP32 only in the last j slices before the recompute
  chunk32    the stack's chunk products B_l ... B_1 U in FP32 feeding the FP64 UDT
             (G_scratch from that stack against the FP64 G_scratch)

Conventions follow the reference: B_l = e^{-dtau T} e^{V_l} (stack.jl:319-367), the rank-1 update
of update_greens! (fields.jl:271-286) on G = [I + B_{l-1} .. B_1 B_M .. B_l]^-1 before the
wrap with B_l (local_updates.jl:7-14, stack.jl:594-603), the density Hirsch field of the
attractive model (fields.jl:380-393).

Usage: python oracle/precision_study.py L [thermalisation sweeps]   (config C couplings:
U = 4 attractive, mu = -0.5, beta = 12, dtau = 0.1; L = 24 is config C itself)
"""
from __future__ import annotations

import math
import os
import sys

import numpy as np
import scipy.linalg as sl

HERE = os.path.dirname(os.path.abspath(__file__))
# run as a script, the oracle directory itself is first on the path: use the repository root
sys.path = [q for q in sys.path if os.path.abspath(q or ".") != HERE]
sys.path.insert(0, os.path.dirname(HERE))


def thermalised_field(L: int, sweeps: int, U=4.0, mu=-0.5, beta=12.0, seed=3):
    from oracle import oracle as O
    from tests.conftest import load_pkg, make_chain
    pkg = load_pkg()
    ch, model, par, field = make_chain(O.OracleChain, pkg, L, U, beta, mu=mu, seed=seed)
    ch.set_hopping(model.hopping_matrix())
    ch.set_field(field.conf)
    ch.set_seed(seed)
    ch.build_stack()
    if sweeps:
        ch.sweep(sweeps)
    conf = ch.get_field()
    ch.close()
    return model.hopping_matrix(), conf, par


class Study:
    def __init__(self, T, conf, par, alpha, safe_mult=10, seed=5):
        self.n, self.M = conf.shape
        self.s = safe_mult
        self.dtau = par.delta_tau
        self.alpha = alpha
        self.eT, self.eTi = sl.expm(-self.dtau * T), sl.expm(self.dtau * T)
        self.conf = conf.astype(np.int8)
        self.rng = np.random.default_rng(seed)

    def B(self, l, c):
        return self.eT * np.exp(self.alpha * c[:, l - 1])[None, :]

    def Binv(self, l, c):
        return np.exp(-self.alpha * c[:, l - 1])[:, None] * self.eTi

    @staticmethod
    def udt(X):
        Q, R, P = sl.qr(X, pivoting=True)
        D = np.abs(np.diag(R))
        D[D == 0] = 1
        T = np.empty_like(R)
        T[:, P] = R / D[:, None]
        return Q, D, T

    def greens(self, l, c, chunk32=False):
        """G(l) = [I + B_l .. B_1 B_M .. B_{l+1}]^-1, a UDT every safe_mult slices."""
        n = self.n
        order = list(range(l + 1, self.M + 1)) + list(range(1, l + 1))
        U, D, T = np.eye(n), np.ones(n), np.eye(n)
        cur = U.copy()
        for k, s in enumerate(order):
            if chunk32:
                cur = (self.B(s, c).astype(np.float32) @ cur.astype(np.float32)).astype(np.float64)
            else:
                cur = self.B(s, c) @ cur
            if (k + 1) % self.s == 0 or k == len(order) - 1:
                U, D2, T2 = self.udt(cur * D[None, :])
                T, D, cur = T2 @ T, D2, U.copy()
        Db, Ds = np.maximum(D, 1), np.minimum(D, 1)
        Ti = np.linalg.inv(T)
        mid = (U.T @ Ti) / Db[:, None] + np.diag(Ds)
        return Ti @ np.linalg.solve(mid, U.T / Db[:, None])

    def interval(self, l0, mode, tail=0):
        """max|G_propagated - G_scratch| after the slices l0+1 .. l0+safe_mult."""
        c = self.conf.copy()
        G = self.greens(l0, c)
        a = self.alpha
        f32 = lambda x: x.astype(np.float32)
        for l in range(l0 + 1, l0 + self.s + 1):
            low = mode == "tail" and l > l0 + self.s - tail
            if low:
                G = f32(G).astype(np.float64)
            for i in range(self.n):
                x = c[i, l - 1]
                d = math.exp(-2 * a * x) - 1
                R = 1 + d * (1 - G[i, i])
                p = math.exp(2 * a * x) * R * R
                if p > 1 or self.rng.random() < p:
                    IG = -G[:, i].copy()
                    IG[i] += 1
                    g = (d / R) * G[i, :]
                    if mode == "upd32" or low:
                        G -= np.outer(f32(IG), f32(g)).astype(np.float64)
                    else:
                        G -= np.outer(IG, g)
                    c[i, l - 1] = -x
            if mode == "wrap32" or low:
                G = (f32(self.B(l, c)) @ f32(G) @ f32(self.Binv(l, c))).astype(np.float64)
            else:
                G = self.B(l, c) @ G @ self.Binv(l, c)
        return float(np.abs(G - self.greens(l0 + self.s, c)).max())


def main():
    L = int(sys.argv[1]) if len(sys.argv) > 1 else 8
    sweeps = int(sys.argv[2]) if len(sys.argv) > 2 else 40
    T, conf, par = thermalised_field(L, sweeps)
    alpha = math.acosh(math.exp(0.5 * par.delta_tau * 4.0))
    st = Study(T, conf, par, alpha)
    starts = [0, st.M // 2 - st.M // 2 % st.s, st.M - 2 * st.s]
    print(f"L={L} N={st.n} M={st.M}, field after {sweeps} oracle sweeps; max|G| = "
          f"{np.abs(st.greens(0, st.conf)).max():.3g}")
    for mode, tail in (("fp64", 0), ("upd32", 0), ("wrap32", 0), ("tail", 1), ("tail", 2), ("tail", 3)):
        r = [st.interval(l0, mode, tail) for l0 in starts]
        name = f"tail {tail}" if mode == "tail" else mode
        print(f"  {name:8s} max|dG| after {st.s} slices: " + ", ".join(f"{x:.2e}" for x in r), flush=True)
    for l in starts[:2]:
        d = np.abs(st.greens(l, st.conf, chunk32=True) - st.greens(l, st.conf)).max()
        print(f"  chunk32  slice {l}: max|G_scratch(FP32 chunks) - G_scratch(FP64)| = {d:.2e}", flush=True)


if __name__ == "__main__":
    main()
