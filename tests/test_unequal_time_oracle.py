"""The unequal-time oracle pinned on the CPU.

Checks of the reference's own tests (test/DQMC/unequal_time_stack.jl) and of identities the
mathematics fixes:
  * _find_range_with_value on beta = 2.3, dtau = 0.1, safe_mult = 10          (:1-14)
  * G(k, k) = calculate_greens(k) for every k (HubbardModel(6, 1), beta 15)   (:88-95)
  * G(k, 0) = -G(k, M) (antiperiodicity in tau)                              (:97-102)
  * G(k <- l) against dense slice products, for k >= l and k < l:
        k >= l:  G(k, l) = B(k, l) G_l                 B(k, l) = B_k .. B_{l+1}
        k <  l:  G(k, l) = -(I - G_k) B(l, k)^-1 = -G_k B_k .. B_1 B_M .. B_{l+1}
    with G_l = [I + B_l .. B_1 B_M .. B_{l+1}]^-1 from the C oracle; the reference's
    BigFloat case is HubbardModel(6, 1), beta 5, safe_mult 10, (k, l) = (37, 14)  (:252-300)
"""
import numpy as np
import pytest

from conftest import make_chain, rel_maxdiff


def _chain(pkg, oracle, L, dims, U, beta, safe_mult, seed=3):
    model = pkg.HubbardModel(L, dims, U=U)
    par = pkg.DQMCParameters.create(beta=beta, delta_tau=0.1, safe_mult=safe_mult)
    field = model.choose_field()(par, model)
    field.rand(seed)
    o = oracle.OracleChain(len(model.l), field.flavors, par.slices, safe_mult, par.delta_tau,
                           field.alpha)
    o.set_hopping(model.hopping_matrix())
    o.set_field(field.conf)
    o.build_stack()
    return o, model, par, field


def _B(o, field, b, l):
    _, _, eT2m, _ = o.get_hopping_exp()
    s = 1.0 if b == 0 else -1.0
    return eT2m @ np.diag(np.exp(s * field.alpha * field.conf[:, l - 1]))


def _prod(o, field, b, hi, lo):
    """B_hi .. B_{lo+1} (identity when hi == lo)."""
    n = field.conf.shape[0]
    A = np.eye(n)
    for s in range(lo + 1, hi + 1):
        A = _B(o, field, b, s) @ A
    return A


def _inv_prod(o, field, b, hi, lo):
    """(B_hi .. B_{lo+1})^-1 = B_{lo+1}^-1 .. B_hi^-1 from the inverse slice factors."""
    _, _, _, eT2p = o.get_hopping_exp()
    s = 1.0 if b == 0 else -1.0
    n = field.conf.shape[0]
    A = np.eye(n)
    for l in range(lo + 1, hi + 1):
        A = A @ (np.diag(np.exp(-s * field.alpha * field.conf[:, l - 1])) @ eT2p)
    return A


def test_find_range(pkg, oracle):
    from oracle.unequal_time import UnequalTime
    par = pkg.DQMCParameters.create(beta=2.3, delta_tau=0.1, safe_mult=10)
    assert par.slices == 23
    conf = np.ones((4, 23), dtype=np.int8)
    ut = UnequalTime(np.eye(4), np.eye(4), conf, 0.1, 1.0, 10)
    assert ut.find_range(-81273) == 0 and ut.find_range(0) == 0
    for i in range(1, 24):
        a, b = ut.ranges[ut.find_range(i) - 1]
        assert a <= i <= b
    assert ut.find_range(24) == ut.K + 1 and ut.find_range(1239874) == ut.K + 1


def test_equal_time_and_antiperiodicity(pkg, oracle):
    from oracle.unequal_time import UnequalTime
    o, model, par, field = _chain(pkg, oracle, 6, 1, 1.0, 15.0, 5)
    eTmh, eTph, eT2m, eT2p = o.get_hopping_exp()
    M = par.slices
    ut = UnequalTime(eT2m, eT2p, field.conf, field.alpha, 1.0, 5)
    for k in range(0, M + 1, 7):
        G1 = o.calculate_greens(k)[0]
        G2 = ut.greens(k, k)
        assert np.abs(G1 - G2).max() < 1e-13, k
    for k in range(0, M, 11):
        Gk0 = eTph @ ut.greens(k, 0) @ eTmh
        GkM = eTph @ ut.greens(k, M) @ eTmh
        assert np.abs(Gk0 + GkM).max() < 1e-13, k


@pytest.mark.parametrize("L,dims,U,beta,safe_mult,pairs", [
    (6, 1, 1.0, 5.0, 10, [(37, 14), (14, 37), (50, 0), (0, 50), (10, 10), (21, 20), (20, 21),
                          (11, 10), (10, 11), (40, 1), (1, 40), (30, 30), (50, 49)]),
    (2, 2, -4.0, 2.0, 5, [(13, 4), (4, 13), (20, 0), (0, 20), (5, 5), (6, 5), (5, 6)]),
    (3, 2, 4.0, 3.0, 7, [(29, 3), (3, 29), (15, 14), (14, 15), (30, 1)]),
])
def test_against_dense_products(pkg, oracle, L, dims, U, beta, safe_mult, pairs):
    from oracle.unequal_time import UnequalTime
    o, model, par, field = _chain(pkg, oracle, L, dims, U, beta, safe_mult)
    _, _, eT2m, eT2p = o.get_hopping_exp()
    n = field.conf.shape[0]
    M = par.slices
    for b in range(field.flavors):
        ut = UnequalTime(eT2m, eT2p, field.conf, field.alpha, 1.0 if b == 0 else -1.0, safe_mult)
        for k, l in pairs:
            Gl = o.calculate_greens(l)[b]
            Gk = o.calculate_greens(k)[b]
            # two dense forms of each identity; the dense products lose digits with the
            # condition number of the span, so the bound grows with it (a wrong sign, index
            # or factor is an O(1) error at every size)
            if k >= l:
                A = _prod(o, field, b, k, l)
                refs = (A @ Gl, Gk @ A)
            else:
                A = _prod(o, field, b, l, k)
                refs = (-(np.eye(n) - Gk) @ _inv_prod(o, field, b, l, k),
                        -Gk @ _prod(o, field, b, k, 0) @ _prod(o, field, b, M, l))
            G = ut.greens(k, l)
            err = min(rel_maxdiff(G, r) for r in refs)
            tol = max(1e-10, 1e-20 * np.linalg.cond(A))
            assert err < tol, (b, k, l, err, tol)
            if (L, dims, beta, k, l) == (6, 1, 5.0, 37, 14):
                assert err < 1e-11
    assert M >= max(max(p) for p in pairs)


def test_stack_blocks_are_the_products(pkg, oracle):
    """forward / backward / inverse UDT blocks equal their slice products (:178-245)."""
    from oracle.unequal_time import UnequalTime
    o, model, par, field = _chain(pkg, oracle, 6, 1, 1.0, 5.0, 10)
    _, _, eT2m, eT2p = o.get_hopping_exp()
    ut = UnequalTime(eT2m, eT2p, field.conf, field.alpha, 1.0, 10)
    M = par.slices
    for idx in range(ut.K + 1):
        upto = ut.ranges[idx - 1][1] if idx > 0 else 0
        F = ut.fu[idx] @ np.diag(ut.fd[idx]) @ ut.ft[idx]
        assert rel_maxdiff(F, _prod(o, field, 0, upto, 0)) < 1e-12
        Bk = ut.bu[idx] @ np.diag(ut.bd[idx]) @ ut.bt[idx]
        assert rel_maxdiff(Bk.T, _prod(o, field, 0, M, upto)) < 1e-12
    for idx in range(ut.K):
        a, b = ut.ranges[idx]
        I = ut.iu[idx] @ np.diag(ut.id[idx]) @ ut.it[idx]
        assert rel_maxdiff(I, _inv_prod(o, field, 0, b, a - 1)) < 1e-11
    for low, high in [(14, 37), (0, 50), (3, 7), (10, 20), (9, 31), (25, 25)]:
        U, D, T = ut.inverse_block(low, high)
        ref = _inv_prod(o, field, 0, high, low)
        assert rel_maxdiff(U @ np.diag(D) @ T, ref) < 1e-11, (low, high)


@pytest.mark.parametrize("recalc_mult,tol", [(1, 1e-13), (4, 1e-10)])
def test_combined_iterator_matches_direct(pkg, oracle, recalc_mult, tol):
    """CombinedGreensIterator against greens(mc, k, 0), greens(mc, 0, k) and the equal-time
    greens(mc, k) (unequal_time_stack.jl:115-170: HubbardModel(6, 1), beta 15, safe_mult 5;
    recalculate = safe_mult within 2e-14, 4 safe_mult within 1e-10)."""
    from oracle.unequal_time import UnequalTimeChain, combined_greens
    o, model, par, field = _chain(pkg, oracle, 6, 1, 1.0, 15.0, 5)
    eTmh, eTph, _, _ = o.get_hopping_exp()
    M = par.slices
    ut = UnequalTimeChain(o)
    trip = combined_greens(o, 0, M, recalc_mult * 5, G00=o.calculate_greens(0))
    assert len(trip) == M + 1
    for k, (G0k, Gk0, Gkk) in enumerate(trip):
        assert np.abs(Gk0 - ut.greens(k, 0)).max() < tol, k
        assert np.abs(G0k - ut.greens(0, k)).max() < tol, k
        Geq = eTph @ o.calculate_greens(k)[0] @ eTmh
        assert np.abs(Gkk[0] - Geq).max() < tol, k


@pytest.mark.parametrize("start,stop", [(1, 40), (7, 33), (0, 5)])
def test_combined_iterator_ranges(pkg, oracle, start, stop):
    """start in (0, 1) begins from G(0, 0), any other start from a recalculation; two flavour
    blocks.  Between recalculations the iterator advances by slice products, whose rounding
    grows with their condition: at U = -4 and the default recalculate = 2 safe_mult it
    agrees with the direct G to about 4e-8 after 19 slices (inside the 1e-6 of the
    reference's verify(), greens_iterators.jl:13), with recalculate = safe_mult to 1e-11."""
    from oracle.unequal_time import UnequalTimeChain, combined_greens
    o, model, par, field = _chain(pkg, oracle, 2, 2, -4.0, 4.0, 10)
    ut = UnequalTimeChain(o)
    for rec, tol in ((10, 1e-10), (20, 1e-6)):
        trip = combined_greens(o, start, stop, rec)
        assert len(trip) == stop - start + 1
        for i, (G0l, Gl0, Gll) in enumerate(trip):
            l = start + i
            assert rel_maxdiff(Gl0, ut.greens(l, 0)) < tol, (rec, l)
            assert rel_maxdiff(G0l, ut.greens(0, l)) < tol, (rec, l)
            assert rel_maxdiff(Gll, ut.greens(l, l)) < tol, (rec, l)


def _u0_chain(pkg, oracle, model, beta, field_cls, safe_mult=5):
    par = pkg.DQMCParameters.create(beta=beta, delta_tau=0.1, safe_mult=safe_mult)
    field = field_cls(par, model)
    field.rand(1)
    o = oracle.OracleChain(len(model.l), field.flavors, par.slices, safe_mult, par.delta_tau,
                           field.alpha)
    o.set_hopping(model.hopping_matrix())
    o.set_field(field.conf)
    o.build_stack()
    return o, par


def _ed_density(ed, l, observable, step):
    """sum over EachSitePairByDistance of the integrated <O_trg(tau) O_src(0)> by ED
    (test/ED/ED_tests.jl:300-345), divided by the number of sites."""
    N, B = l.n_bravais, l.n_basis
    d2st = l.bravais_dir2srctrg()
    ref = np.zeros((N, B, B))
    for d in range(N):
        for b1 in range(B):
            for b2 in range(B):
                for s in range(N):
                    src, trg = s + b1 * N, d2st[d][s] + b2 * N
                    if observable == "charge":
                        v = ed.integrated(ed.number(trg), ed.number(src), step)
                    else:
                        (A, f), (C, _) = ed.spin(trg, observable), ed.spin(src, observable)
                        v = (f * f * ed.integrated(A, C, step)).real
                    ref[d, b1, b2] += v
    return ref / len(l)


@pytest.mark.parametrize("lat", ["honeycomb", "square"])
@pytest.mark.parametrize("beta", [1.0, 8.9])
@pytest.mark.parametrize("magnetic", [False, True])
def test_density_susceptibilities_match_ed_at_u0(pkg, oracle, lat, beta, magnetic):
    """charge_density_susceptibility and spin_density_susceptibility(x, y, z) against exact
    diagonalisation at U = 0, where DQMC is exact (test/ED/ED_tests.jl:100-146, 300-345:
    HubbardModel(Honeycomb(2, 1), U = 0), beta in (1.0, 8.9), atol = rtol = 1e-13); one
    (density) and two (magnetic) flavour blocks.  The reduced charge kernel is the full one
    minus beta <n_src><n_trg>."""
    from oracle.ed import HubbardED
    from oracle.unequal_time import density_susceptibility
    l = pkg.Honeycomb(2, 1) if lat == "honeycomb" else pkg.SquareLattice(2, 2)
    model = pkg.HubbardModel(l=l, U=0.0)
    o, par = _u0_chain(pkg, oracle, model, beta,
                       pkg.MagneticHirschField if magnetic else pkg.DensityHirschField)
    eTmh, eTph, _, _ = o.get_hopping_exp()
    Gcur = np.stack([eTph @ G @ eTmh for G in o.calculate_greens(0)])
    ed = HubbardED(model, beta)
    for obs in ("charge", "x", "y", "z"):
        ref = _ed_density(ed, l, obs, par.delta_tau)
        got = density_susceptibility(o, l, Gcur, obs, recalculate=10)
        assert np.allclose(got, ref, atol=1e-12, rtol=1e-12), (obs, np.abs(got - ref).max())
    N, B = l.n_bravais, l.n_basis
    d2st = l.bravais_dir2srctrg()
    occ = np.array([ed.expect(ed.number(k)) for k in range(len(l))])
    disc = np.zeros((N, B, B))
    for d in range(N):
        for b1 in range(B):
            for b2 in range(B):
                s = np.arange(N)
                disc[d, b1, b2] = beta * (occ[s + b1 * N] * occ[d2st[d] + b2 * N]).sum()
    red = density_susceptibility(o, l, Gcur, "charge", recalculate=10, reduced=True)
    full = _ed_density(ed, l, "charge", par.delta_tau)
    assert np.allclose(red, full - disc / len(l), atol=1e-12, rtol=1e-12)
