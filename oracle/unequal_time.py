"""Unequal-time Green's functions on the CPU: restatement of the reference's UnequalTimeStack.

TEST INFRASTRUCTURE ONLY (the checker of dqmc_unequal_time_greens); never loaded by the
product path.

Follows src/flavors/DQMC/unequal_time_stack.jl:
  build_stack(mc, ::UnequalTimeStack)      :124-177  forward / backward / inverse UDT stacks
  _find_range_with_value                    :338-367
  compute_inverse_udt_block!                :383-459  B_{low+1}^-1 .. B_high^-1 = U D T
  compute_forward_udt_block!                :473-494  B_slice .. B_1 = Ul Dl Tl
  compute_backward_udt_block!               :508-531  B_M .. B_{slice+1} = (Ur Dr Tr)^T
  calculate_greens_full1!                   :535-617  G(k <- l), k >= l
  calculate_greens_full2!                   :620-697  G(k <- l), k <  l
  greens(mc, k, l) -> _greens!              greens.jl:114-125 (the physical form)
with the slice matrices of stack.jl:319-367 and the reference's own decompositions
(udt_AVX_pivot!, rdivp! from the C oracle, dqmc_oracle.c).  Matrix products are numpy's.
Each flavour block is handled independently (BlockDiagonal, blockdiagonal.jl).
"""
from __future__ import annotations

import numpy as np

from . import oracle as O


def vmaxinv(v):
    return 1.0 / np.maximum(1.0, v)


def vmin(v):
    return np.minimum(1.0, v)


class UnequalTime:
    """The unequal-time stack of one flavour block for a fixed field.

    eT2m = e^{-dtau T}, eT2p = e^{+dtau T}; ev(l) = the interaction diagonal of slice l
    (1-based), exp(power alpha x) with the block's sign; ranges = generate_chunks(M, safe_mult).
    """

    def __init__(self, eT2m, eT2p, conf, alpha, sign, safe_mult):
        self.eT2m, self.eT2p = eT2m, eT2p
        self.conf = np.asarray(conf)
        self.alpha, self.sign = alpha, sign
        self.n, self.M = self.conf.shape
        self.ranges = O.generate_chunks(self.M, safe_mult)
        self.K = len(self.ranges)
        self.build()

    # ---- slice matrices (stack.jl:319-367) ---------------------------------------------
    def ev(self, l, power=1.0):
        return np.exp(power * self.sign * self.alpha * self.conf[:, l - 1])

    def left(self, l, M):              # eT2 eV M
        return self.eT2m @ (self.ev(l)[:, None] * M)

    def dagger_left(self, l, M):       # eV^T eT2^T M
        return self.ev(l)[:, None] * (self.eT2m.T @ M)

    def inv_left(self, l, M):          # eV^-1 eT2^-1 M
        return self.ev(l, -1.0)[:, None] * (self.eT2p @ M)

    def inv_right(self, l, M):         # M eV^-1 eT2^-1
        return (M * self.ev(l, -1.0)[None, :]) @ self.eT2p

    # ---- the stack (unequal_time_stack.jl:124-177) ---------------------------------------
    def build(self):
        n, K = self.n, self.K
        I = np.eye(n)
        self.fu, self.fd, self.ft = [I] * (K + 1), [np.ones(n)] * (K + 1), [I] * (K + 1)
        for idx in range(1, K + 1):
            cur = self.fu[idx - 1]
            r0, r1 = self.ranges[idx - 1]
            for s in range(r0, r1 + 1):
                cur = self.left(s, cur)
            U, D, T, _ = O.udt(cur * self.fd[idx - 1][None, :])
            self.fu[idx], self.fd[idx], self.ft[idx] = U, D, T @ self.ft[idx - 1]
        self.bu, self.bd, self.bt = [I] * (K + 1), [np.ones(n)] * (K + 1), [I] * (K + 1)
        for idx in range(K, 0, -1):
            cur = self.bu[idx]
            r0, r1 = self.ranges[idx - 1]
            for s in range(r1, r0 - 1, -1):
                cur = self.dagger_left(s, cur)
            U, D, T, _ = O.udt(cur * self.bd[idx][None, :])
            self.bu[idx - 1], self.bd[idx - 1], self.bt[idx - 1] = U, D, T @ self.bt[idx]
        self.iu, self.id, self.it = [None] * K, [None] * K, [None] * K
        for idx in range(1, K + 1):
            cur = I
            r0, r1 = self.ranges[idx - 1]
            for s in range(r1, r0 - 1, -1):
                cur = self.inv_left(s, cur)
            U, D, T, _ = O.udt(cur)
            self.iu[idx - 1], self.id[idx - 1], self.it[idx - 1] = U, D, T

    def find_range(self, val):
        """_find_range_with_value (:338-367): 1-based index of the range holding val, 0 below
        all ranges, K + 1 above."""
        if val < 1:
            return 0
        if val > self.ranges[-1][1]:
            return self.K + 1
        for i, (a, b) in enumerate(self.ranges, start=1):
            if a <= val <= b:
                return i
        raise ValueError(val)

    def inverse_block(self, low, high):
        """B_{low+1}^-1 .. B_high^-1 = U D T (:383-459)."""
        n = self.n
        lower = self.find_range(low) + 1
        upper = self.find_range(high + 1) - 1
        U, D, T = np.eye(n), np.ones(n), np.eye(n)
        for idx in range(lower, upper + 1):
            tmp1 = (D[:, None] * (T @ self.iu[idx - 1])) * self.id[idx - 1][None, :]
            tmp2, D, tmp1, _ = O.udt(tmp1)
            T = tmp1 @ self.it[idx - 1]
            U = U @ tmp2
        lower_slice = self.ranges[lower - 1][0] if lower <= self.K else self.ranges[-1][1] + 1
        upper_slice = self.ranges[upper - 1][1] if upper > 0 else 0
        top = min(lower_slice - 1, high)
        for s in range(top, low, -1):
            U = self.inv_left(s, U)
        if top >= low + 1:
            U, D, tmp1, _ = O.udt(U * D[None, :])
            T = tmp1 @ T
        for s in range(max(upper_slice + 1, top + 1), high + 1):
            T = self.inv_right(s, T)
        return U, D, T

    def forward_block(self, slice_):
        """B_slice .. B_1 = Ul Dl Tl (:473-494)."""
        idx = max(0, self.find_range(slice_) - 1)
        T = self.fu[idx]
        target = self.ranges[idx - 1][1] + 1 if idx > 0 else 1
        for l in range(target, slice_ + 1):
            T = self.left(l, T)
        U, D, tmp, _ = O.udt(T * self.fd[idx][None, :])
        return U, D, tmp @ self.ft[idx]

    def backward_block(self, slice_):
        """B_M .. B_{slice+1} = (Ur Dr Tr)^T (:508-531)."""
        idx = self.find_range(slice_) + 1
        U = self.bu[idx - 1]
        target = self.ranges[idx - 1][0] - 1 if idx <= self.K else self.ranges[-1][1]
        for l in range(target, slice_, -1):
            U = self.dagger_left(l, U)
        U, D, tmp, _ = O.udt(U * self.bd[idx - 1][None, :])
        return U, D, tmp @ self.bt[idx - 1]

    def greens(self, k, l):
        """The effective G(k <- l) (calculate_greens, :312-326)."""
        return self._full1(k, l) if k >= l else self._full2(k, l)

    def _full1(self, k, l):
        U, D, T = self.inverse_block(l, k)
        Ul, Dl, Tl = self.forward_block(l)
        Ur, Dr, Tr = self.backward_block(k)
        G = Dl[:, None] * ((Tl @ Tr.T) * Dr[None, :])
        Tr, Dr, G, piv = O.udt(G, False)
        Tl = Ul @ Tr
        Ur = O.rdivp(Ur, G, piv)
        Tr = (vmaxinv(D)[:, None] * (U.T @ Tl)) * vmin(Dr)[None, :]
        Tl = (vmin(D)[:, None] * (T @ Ur)) * vmaxinv(Dr)[None, :]
        Tl = Tl + Tr
        Tr, Dl, Tl, piv = O.udt(Tl, False)
        Ul = O.rdivp(np.diag(vmaxinv(Dr)), Tl, piv)
        Ul = (Ul * (1.0 / Dl)[None, :]) @ Tr.T
        G = Ul * vmaxinv(D)[None, :]
        return Ur @ (G @ U.T)

    def _full2(self, k, l):
        U, D, T = self.inverse_block(k, l)
        Ul, Dl, Tl = self.forward_block(k)
        Ur, Dr, Tr = self.backward_block(l)
        G = (Dl[:, None] * (Tl @ Tr.T)) * Dr[None, :]
        Tr, Dr, G, piv = O.udt(G, False)
        Tl = Ul @ Tr
        Ul = (vmaxinv(D)[:, None] * (U.T @ Tl)) * vmin(Dr)[None, :]
        Us = O.rdivp(T @ Ur, G, piv)
        Tr = (vmin(D)[:, None] * Us) * vmaxinv(Dr)[None, :]
        Tr = Tr + Ul
        Ul, Dl, Tr, piv = O.udt(Tr, False)
        Us = O.rdivp(np.diag(vmin(Dr)), Tr, piv)
        Ur = ((Us * (1.0 / Dl)[None, :]) @ Ul.T) * vmin(D)[None, :]
        return -(Tl @ (Ur @ T))


class UnequalTimeChain:
    """The unequal-time stacks of every flavour block of an oracle chain's current field."""

    def __init__(self, chain, safe_mult=None):
        self.eTmh, self.eTph, eT2m, eT2p = chain.get_hopping_exp()
        conf = chain.get_field()
        sm = chain.safe_mult if safe_mult is None else safe_mult
        self.blocks = [UnequalTime(eT2m, eT2p, conf, chain.alpha, 1.0 if b == 0 else -1.0, sm)
                       for b in range(chain.flv)]

    def greens(self, k, l, physical=True):
        """greens(mc, k, l) (physical) or calculate_greens(mc, k, l): (flavours, n, n)."""
        out = [ut.greens(k, l) for ut in self.blocks]
        if physical:
            out = [self.eTph @ G @ self.eTmh for G in out]
        return np.stack(out)


def unequal_time_greens(chain, k, l, physical=True, safe_mult=None):
    """greens(mc, k, l) of an oracle chain's current field (flavours, n, n)."""
    return UnequalTimeChain(chain, safe_mult).greens(k, l, physical)


def combined_greens(chain, start=0, stop=None, recalculate=None, G00=None, safe_mult=None):
    """CombinedGreensIterator(mc; start, stop, recalculate) iterated to the end
    (measurements/greens_iterators.jl:181-407): for l = start..stop the physical triple
    (G(0, l), G(l, 0), G(l, l)), each (flavours, n, n).  G00: the chain's effective G when its
    current slice is 1 (the iterator takes it instead of calculate_greens_full1!(0, 0));
    recalculate defaults to 2 safe_mult."""
    utc = UnequalTimeChain(chain, safe_mult)
    sm = chain.safe_mult if safe_mult is None else safe_mult
    M = chain.M
    stop = M if stop is None else stop
    rec = 2 * sm if recalculate is None else recalculate
    per_block = [_combined_block(ut, utc.eTph, utc.eTmh, None if G00 is None else G00[b],
                                 start, stop, rec, sm)
                 for b, ut in enumerate(utc.blocks)]
    return [tuple(np.stack([blk[i][j] for blk in per_block]) for j in range(3))
            for i in range(len(per_block[0]))]


def _combined_block(u, eTph, eTmh, G00, start, stop, rec, sm):
    n = u.n
    I = np.eye(n)

    def phys(X):
        return eTph @ X @ eTmh

    out = []

    def recalc_at(l):
        Gl0, G0l, Gll = u.greens(l, 0), u.greens(0, l), u.greens(l, l)
        trip = (phys(G0l), phys(Gl0), phys(Gll))
        uU, uD, uT, _ = O.udt(Gll.copy())
        Ul, Dl, Tl, _ = O.udt(Gl0.copy())
        Ur, Dr, Tr, _ = O.udt(G0l.copy())
        return trip, [Ul, Dl, Tl, Ur, Dr, Tr, uU, uD, uT]

    if start in (0, 1):
        G = u.greens(0, 0) if G00 is None else G00
        Ul, Dl, Tl, _ = O.udt(G.copy())
        uU, uD, uT = Ul.copy(), Dl.copy(), Tl.copy()
        Ur, Dr, Tr, _ = O.udt(G - I)
        st = [Ul, Dl, Tl, Ur, Dr, Tr, uU, uD, uT]
        if start == 0:
            P = phys(G)
            out.append((P, P.copy(), P.copy()))
        l = 1
    else:
        trip, st = recalc_at(start)
        out.append(trip)
        l = start + 1
    shift = start if start != 1 else 0
    while l <= stop:
        Ul, Dl, Tl, Ur, Dr, Tr, uU, uD, uT = st
        if (l - shift) % rec == 0:
            trip, st = recalc_at(l)
            out.append(trip)
            l += 1
            continue
        Ul = u.left(l, Ul)
        Tr = u.inv_right(l, Tr)
        uU = u.left(l, uU)
        uT = u.inv_right(l, uT)
        if ((l - shift) % rec) % sm == 0:
            # stabilisation: the three products are returned, then re-decomposed
            X = Ul * Dl[None, :]
            Gl0 = X @ Tl
            Ul, Dl, X, _ = O.udt(X)
            Tl = X @ Tl
            X = Dr[:, None] * Tr
            G0l = Ur @ X
            Y, Dr, Tr, _ = O.udt(X)
            Ur = Ur @ Y
            X = uU * uD[None, :]
            Gll = X @ uT
            C, uD, X, _ = O.udt(X)
            uU = X @ uT
            uT = uD[:, None] * uU
            X, uD, uT, _ = O.udt(uT)
            uU = C @ X
        else:
            Gl0 = (Ul * Dl[None, :]) @ Tl
            G0l = (Ur * Dr[None, :]) @ Tr
            Gll = (uU * uD[None, :]) @ uT
        out.append((phys(G0l), phys(Gl0), phys(Gll)))
        st = [Ul, Dl, Tl, Ur, Dr, Tr, uU, uD, uT]
        l += 1
    return out


def density_susceptibility(chain, lattice, G_current, observable="charge", recalculate=None,
                           reduced=False, G00=None, total_flavors=2):
    """charge_density_susceptibility (observable "charge") or spin_density_susceptibility
    (observable "x", "y", "z") for the chain's current field, once: apply!(::TimeIntegral)
    (measurements/generic.jl:337-370: l = 0..M with trapezoid weights dtau, 1/2 at l = 0 and
    l = M, G00 = the chain's current physical G), the EachSitePairByDistance accumulation
    (generic.jl:434-460: temp[dir, b1, b2] over Bravais src, trg = dir2srctrg[dir][src]),
    the kernels full/reduced_cdc_kernel (constructors/charge_density.jl:62-170, flavour pairs
    (f1, f2)) and full/reduced_sdc_{x,y,z}_kernel (constructors/spin_density.jl:68-330, one
    call), DiagonallyRepeating for one unique flavour (with total_flavors) and BlockDiagonal
    otherwise, and finish! (generic.jl:574-586: temp / number of sites).
    G_current: (flavours, n, n) physical.  Returns (Ndir, B, B)."""
    M, dtau, flv = chain.M, chain.delta_tau, chain.flv
    trip = combined_greens(chain, 0, M, recalculate, G00)
    d2st = lattice.bravais_dir2srctrg()
    N, B = lattice.n_bravais, lattice.n_basis
    out = np.zeros((N, B, B))
    src = np.arange(N)
    G00p = G_current
    for l, (G0l, Gl0, Gll) in enumerate(trip):
        w = (0.5 if l in (0, M) else 1.0) * dtau
        for d in range(N):
            trg = d2st[d]
            for b2 in range(B):
                for b1 in range(B):
                    i = src + b1 * N
                    j = trg + b2 * N
                    idm = ((i == j) & (l == 0)).astype(float)

                    def conn(f1, f2):
                        return (idm - G0l[f1][j, i]) * Gl0[f2][i, j]

                    def occ(G, f, k):
                        return 1 - G[f][k, k]

                    if flv == 1:
                        if observable == "charge":
                            tf = total_flavors
                            val = tf * conn(0, 0)
                            if not reduced:
                                val = val + tf * tf * occ(Gll, 0, i) * occ(G00p, 0, j)
                        else:
                            val = 2 * conn(0, 0)
                    elif observable == "charge":
                        val = np.zeros(N)
                        for f1 in range(flv):
                            for f2 in range(flv):
                                if not reduced:
                                    val = val + occ(Gll, f1, i) * occ(G00p, f2, j)
                                if f1 == f2:
                                    val = val + conn(f1, f1)
                    elif observable == "z":
                        val = conn(0, 0) + conn(1, 1)
                        if not reduced:
                            val = val + (occ(Gll, 0, i) * occ(G00p, 0, j) - occ(Gll, 0, i) * occ(G00p, 1, j)
                                         - occ(Gll, 1, i) * occ(G00p, 0, j) + occ(Gll, 1, i) * occ(G00p, 1, j))
                    else:   # x, y: identical for block-diagonal G
                        val = conn(0, 1) + conn(1, 0)
                    out[d, b1, b2] += w * val.sum()
    return out / (N * B)
