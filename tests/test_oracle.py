"""Pins the CPU oracle (no GPU) against the reference's known answers, closed forms,
library routines and brute force -- each check chosen so that a dropped term, wrong sign,
wrong index or transposed operand fails it.
"""
import json
import os

import numpy as np
import pytest
import scipy.linalg as sla

from conftest import ROOT, make_chain, rel_maxdiff

GOLDEN = os.path.join(ROOT, "tests", "golden")


def graded(n, rng, decades):
    q1, _ = np.linalg.qr(rng.standard_normal((n, n)))
    q2, _ = np.linalg.qr(rng.standard_normal((n, n)))
    s = 10.0 ** np.linspace(0, -decades, n)
    rng.shuffle(s)
    return np.asfortranarray(q1 @ np.diag(s) @ q2)


# ---- linear algebra (UDT.jl, real.jl) ------------------------------------------------------
@pytest.mark.parametrize("n", [1, 5, 16, 40])
def test_udt_matches_lapack_geqp3(oracle, n):
    """udt_AVX_pivot! == LAPACK's column-pivoted QR (scipy geqp3) up to column signs:
    same pivots, |diag R| = D, U D T = X, T upper triangular after the pivot (test/linalg.jl:97-124)."""
    rng = np.random.default_rng(n)
    X = graded(n, rng, 6)
    U, D, T, piv = oracle.udt(X, True)
    Q, R, P = sla.qr(X, pivoting=True)
    assert piv.tolist() == P.tolist()
    assert rel_maxdiff(D, np.abs(np.diag(R))) < 1e-12
    assert np.abs(U @ np.diag(D) @ T - X).max() < 1e-13 * np.abs(X).max() * max(1, n / 8)
    assert np.abs(U.T @ U - np.eye(n)).max() < 1e-13
    Tp = T[:, piv]
    assert np.abs(np.tril(Tp, -1)).max() == 0.0
    signs = np.sign(np.diag(R)) * np.sign(np.diag(Tp))
    assert np.abs(U * signs - Q).max() < 1e-11


def test_udt_rank_one_input(oracle):
    """A rank-deficient (kron) input: zero R diagonal entries are replaced by 1 in D (UDT.jl:292-302)."""
    v = np.arange(1.0, 7.0)
    X = np.asfortranarray(np.outer(v, v))
    U, D, T, piv = oracle.udt(X, True)
    assert np.abs(U @ np.diag(D) @ T - X).max() < 1e-12
    assert np.all(D > 0)


def test_udt_dirty_and_rdivp(oracle):
    """Val(false): U D triu(T) P' = X; rdivp!(A, T, pivot) = A P / triu(T) (test/linalg.jl:115-130)."""
    rng = np.random.default_rng(3)
    X = graded(12, rng, 5)
    U, D, T, piv = oracle.udt(X, False)
    P = np.zeros((12, 12))
    P[piv, np.arange(12)] = 1.0
    assert np.abs(U @ np.diag(D) @ np.triu(T) @ P.T - X).max() < 1e-13
    A = rng.standard_normal((12, 12))
    assert rel_maxdiff(oracle.rdivp(A, T, piv), A @ P @ np.linalg.inv(np.triu(T))) < 1e-12


def test_calc_greens_udt_is_the_inverse(oracle):
    """calculate_greens_AVX! = [I + Ul Dl Tl Tr' Dr Ur']^-1 (stack.jl:426-441)."""
    rng = np.random.default_rng(4)
    n = 10
    Ul, Dl, Tl, _ = oracle.udt(graded(n, rng, 3), True)
    Ur, Dr, Tr, _ = oracle.udt(graded(n, rng, 3), True)
    G = oracle.calc_greens_udt(Ul, Dl, Tl, Ur, Dr, Tr)
    ref = np.linalg.inv(np.eye(n) + Ul @ np.diag(Dl) @ Tl @ Tr.T @ np.diag(Dr) @ Ur.T)
    assert rel_maxdiff(G, ref) < 1e-10


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_gemm(oracle, ta, tb):
    rng = np.random.default_rng(ta + 2 * tb)
    for n in (1, 7, 33, 70):
        A, B = rng.standard_normal((n, n)), rng.standard_normal((n, n))
        ref = (A.T if ta else A) @ (B.T if tb else B)
        assert rel_maxdiff(oracle.gemm(A, B, ta, tb), ref) < 1e-13


def test_expm_matches_scipy(pkg, oracle):
    """exp(c T) of the hopping matrix == scipy.linalg.expm (stack.jl:235-239)."""
    T = pkg.HubbardModel(l=pkg.Honeycomb(3), U=1.0, mu=0.4).hopping_matrix()
    for c in (-0.05, 0.05, -0.1, 0.1, -1.0):
        E = oracle.expm_sym(T, c)
        assert rel_maxdiff(E, sla.expm(c * T)) < 1e-13
        assert np.array_equal(E, E.T)


# ---- stack / Green's function (stack.jl) ---------------------------------------------------
def _chain(pkg, oracle, L, U, beta, seed=11, safe_mult=10, mu=0.0):
    o, model, par, field = make_chain(oracle.OracleChain, pkg, L, U, beta, seed=seed,
                                      safe_mult=safe_mult, mu=mu)
    o.set_hopping(model.hopping_matrix())
    o.set_field(field.conf)
    return o, model, par, field


def _slice_B(o, field, flv_b, l):
    """B(l) = eT2 eV(l) for flavour block b (stack.jl:319-327, fields.jl:380-386, 429-438)."""
    _, _, eT2m, _ = o.get_hopping_exp()
    s = 1.0 if flv_b == 0 else -1.0
    return eT2m @ np.diag(np.exp(s * field.alpha * field.conf[:, l - 1]))


def _dense_greens(o, field, par, b, l):
    """G_l = [I + B(l-1)..B(1) B(M)..B(l)]^-1 by dense products (small, well conditioned cases)."""
    n = field.conf.shape[0]
    A = np.eye(n)
    for k in list(range(l, par.slices + 1)):
        A = _slice_B(o, field, b, k) @ A
    for k in range(1, l):
        A = _slice_B(o, field, b, k) @ A
    return np.linalg.inv(np.eye(n) + A)


def test_u0_closed_form(pkg, oracle):
    """U = 0: G = [I + exp(-beta T)]^-1 at 1e-12 (test/flavortests_DQMC.jl:355-386)."""
    for model in (pkg.HubbardModel(7, 2, U=0.0), pkg.HubbardModel(8, 2, U=0.0, mu=1.0)):
        for beta in (1.0, 10.0):
            par = pkg.DQMCParameters.create(beta=beta, delta_tau=0.1, safe_mult=5)
            n = len(model.l)
            o = oracle.OracleChain(n, 1, par.slices, 5, 0.1, 0.0)
            o.set_hopping(model.hopping_matrix())
            o.set_seed(1)
            o.build_stack()
            o.sweep(1)
            w, V = np.linalg.eigh(model.hopping_matrix())
            G = V @ np.diag(1 / (1 + np.exp(-beta * w))) @ V.T
            assert np.abs(o.greens(physical=True)[0] - G).max() < 1e-12 + 1e-12 * np.abs(G).max()


@pytest.mark.parametrize("U", [2.0, -2.0])
def test_calculate_greens_all_slices_vs_dense(pkg, oracle, U):
    """calculate_greens(mc, k) is G_{k+1} = [I + B(k)..B(1) B(M)..B(k+1)]^-1 for every k
    (test/flavortests_DQMC.jl:304-311)."""
    o, model, par, field = _chain(pkg, oracle, 2, U, 1.0, safe_mult=5)
    for k in range(par.slices + 1):
        G = o.calculate_greens(k)
        for b in range(field.flavors):
            ref = _dense_greens(o, field, par, b, k + 1 if k < par.slices else 1)
            assert rel_maxdiff(G[b], ref) < 1e-11, (k, b)


def test_wraps_match_from_scratch(pkg, oracle):
    """Ten consecutive wraps agree with the from-scratch G within 1e-9 (flavortests_DQMC.jl:297-301:
    HubbardModel(8, 1) -- an 8-site chain, U = 1 -- at beta = 5)."""
    model = pkg.HubbardModel(8, 1, U=1.0)
    par = pkg.DQMCParameters.create(beta=5.0, safe_mult=50)
    field = model.choose_field()(par, model)
    field.rand(11)
    o = oracle.OracleChain(8, 1, par.slices, 50, par.delta_tau, field.alpha)
    o.set_hopping(model.hopping_matrix())
    o.set_field(field.conf)
    o.build_stack()
    for _ in range(10):
        o.propagate()
    cs, _, _ = o.state()
    assert rel_maxdiff(o.greens(), o.calculate_greens(cs - 1)) < 1e-9


def test_build_equals_reverse_build(pkg, oracle):
    """build_stack + M propagates == reverse_build_stack + propagate (flavortests_DQMC.jl:313-342)."""
    o1, model, par, field = _chain(pkg, oracle, 4, 4.0, 2.0)
    o1.build_stack_forward()
    for _ in range(par.slices):
        o1.propagate()
    o2, *_ = _chain(pkg, oracle, 4, 4.0, 2.0)
    o2.build_stack()
    assert o1.state() == o2.state()
    assert rel_maxdiff(o1.greens(), o2.greens()) < 1e-10
    for s in range(1, len(par.chunks) + 2):
        for a, b in zip(o1.stack_slot(s), o2.stack_slot(s)):
            assert np.abs(a - b).max() < 1e-10 * max(1.0, np.abs(b).max())


@pytest.mark.parametrize("U", [4.0, -4.0])
def test_detratio_is_the_dense_determinant_ratio(pkg, oracle, U):
    """R_b = 1 + Delta_b (1 - G_ii) equals det(I + A'_l)/det(I + A_l) of the flipped chain
    (fields.jl:58-84, 388-449): brute force at 2x2."""
    o, model, par, field = _chain(pkg, oracle, 2, U, 1.0, safe_mult=5)
    a = field.alpha
    for l in (1, 4, par.slices):
        G = o.calculate_greens(l - 1)
        for i in range(4):
            x = field.conf[i, l - 1]
            dE = -2.0 * a * x
            deltas = [np.exp(dE) - 1.0] if field.flavors == 1 else [np.exp(dE) - 1.0, np.exp(-dE) - 1.0]
            for b, d in enumerate(deltas):
                R = 1.0 + d * (1.0 - G[b][i, i])
                n = 4
                A = np.eye(n)
                for k in list(range(l, par.slices + 1)) + list(range(1, l)):
                    A = _slice_B(o, field, b, k) @ A
                f2 = field.conf.copy()
                f2[i, l - 1] *= -1
                A2 = np.eye(n)
                saved = field.conf.copy()
                field.conf[:, :] = f2
                for k in list(range(l, par.slices + 1)) + list(range(1, l)):
                    A2 = _slice_B(o, field, b, k) @ A2
                field.conf[:, :] = saved
                ratio = np.linalg.det(np.eye(n) + A2) / np.linalg.det(np.eye(n) + A)
                assert abs(R - ratio) < 1e-10 * max(1.0, abs(ratio)), (l, i, b)


def test_sherman_morrison_equals_recompute(pkg, oracle):
    """After the accepted flips of one slice visit, the rank-1 updated G equals the
    from-scratch G of the flipped field (fields.jl:271-286)."""
    o, model, par, field = _chain(pkg, oracle, 4, -4.0, 2.0)
    o.set_seed(5)
    o.build_stack()
    acc = o.sweep_spatial()
    assert acc > 0
    cs, _, _ = o.state()
    o2, *_ = _chain(pkg, oracle, 4, -4.0, 2.0)
    o2.set_field(o.get_field())
    assert rel_maxdiff(o.greens(), o2.calculate_greens(cs - 1)) < 1e-9


def test_config_a_fixture_regression(pkg, oracle):
    """The committed oracle vectors (tests/golden/config_a.npz) are reproduced."""
    d = np.load(os.path.join(GOLDEN, "config_a.npz"))
    model = pkg.HubbardModel(4, 2, U=4.0)
    par = pkg.DQMCParameters.create(beta=4.0, delta_tau=0.1, safe_mult=10)
    o = oracle.OracleChain(16, 1, par.slices, 10, 0.1, float(d["alpha"]))
    o.set_hopping(model.hopping_matrix())
    o.set_field(d["conf"])
    for s in (0, 13, par.slices):
        assert rel_maxdiff(o.calculate_greens(s)[0], d[f"G_slice{s}"]) < 1e-12
    o.set_uniforms(d["uniforms"])
    o.set_log(len(d["uniforms"]))
    o.build_stack()
    o.sweep(1)
    dec, p = o.get_log()
    assert np.array_equal(dec, d["accepted"])
    assert np.array_equal(o.get_field(), d["final_conf"])


def test_state_machine_visits(pkg, oracle):
    """A sweep visits slices 1..M then M..1 and ends at slice 1 moving up (stack.jl:605-730)."""
    o, model, par, field = _chain(pkg, oracle, 2, 1.0, 1.5)
    o.build_stack()
    seen = []
    for _ in range(2 * par.slices):
        seen.append(o.state()[:2])
        o.propagate()
    M = par.slices
    assert [s for s, _ in seen] == list(range(1, M + 1)) + list(range(M, 0, -1))
    assert o.state()[:2] == (1, 1)


# ---- exact diagonalisation (test/ED) --------------------------------------------------------
def test_ed_oracle_free_fermions(pkg):
    """ED at U = 0 equals the free-fermion closed form (test/ED/ED_tests.jl:102-200)."""
    from oracle.ed import HubbardED
    m = pkg.HubbardModel(2, 2, U=0.0, mu=0.3)
    e = HubbardED(m, beta=1.5)
    T = m.hopping_matrix()
    w, V = np.linalg.eigh(T)
    G = V @ np.diag(1 / (1 + np.exp(-1.5 * w))) @ V.T
    Ged = e.greens()
    assert np.abs(Ged[:4, :4] - G).max() < 1e-12 and np.abs(Ged[4:, 4:] - G).max() < 1e-12
    assert np.abs(Ged[:4, 4:]).max() < 1e-14
    assert abs(e.energy() - 2 * np.sum(w / (1 + np.exp(1.5 * w)))) < 1e-12


@pytest.mark.parametrize("case", ["repulsive", "attractive_mu1"])
def test_oracle_dqmc_matches_ed(pkg, oracle, case):
    """The reference's finite-U ED test (test/ED/ED_tests.jl:402-484): 2x2, beta=1, dtau=0.1,
    safe_mult=5, 5k + 5k sweeps; atol 3.05 dtau^2, rtol 2 dtau^2 -- run on the oracle."""
    ed = json.load(open(os.path.join(GOLDEN, "ed_2x2.json")))[case]
    model = pkg.HubbardModel(2, 2, U=ed["U"], mu=ed["mu"], t=1.0)
    mc = pkg.DQMC(model, beta=1.0, delta_tau=0.1, safe_mult=5, thermalization=5000, sweeps=5000,
                  measure_rate=1, seed=77, chain_factory=oracle.OracleChain)
    assert mc.run(verbose=False) == pkg.ExitCode.SUCCESS
    ms = mc.measurements["equal_time"]
    atol, rtol = 3.05 * 0.01, 2 * 0.01
    close = lambda a, b: np.all(np.abs(np.asarray(a) - np.asarray(b)) <= atol + rtol * np.abs(np.asarray(b)))
    assert close(ms.mean("total_energy"), ed["energy"])
    G = ms.mean("greens")
    Ged = np.asarray(ed["greens"])
    nb = G.shape[0]
    for b in range(nb):
        assert close(G[b], Ged[b * 4:(b + 1) * 4, b * 4:(b + 1) * 4])
    assert close(ms.mean("occupation"), 1 - np.diag(Ged)[: 4 * nb])
    assert close(ms.mean("double_occupancy"), ed["double_occupancy"])


@pytest.mark.parametrize("U_ref", [4.0, -4.0])
@pytest.mark.parametrize("move", ["flip", "spatial", "random"])
def test_global_update_probability_is_the_dense_weight_ratio(pkg, oracle, U_ref, move):
    """propose_global_from_conf (global_updates.jl:136-169) against brute force on 2x2,
    beta = 1: p = exp(-dE) prod_b det(I + B_M..B_1)_new / det(..)_old, the per-flavour
    determinants of the dense slice-matrix products (squared for the density field's one
    unique flavour), the way test/updates.jl:186-245 pins global moves."""
    from conftest import make_chain
    o, model, par, field = make_chain(oracle.OracleChain, pkg, 2, U_ref, 1.0, seed=31)
    o.set_hopping(model.hopping_matrix())
    o.set_field(field.conf)
    o.set_seed(5)
    o.build_stack()
    n, M = o.n, o.M
    rng = np.random.default_rng(4)
    c_old = o.get_field()
    if move == "flip":
        c_new = -c_old
    elif move == "spatial":
        c_new = c_old[rng.permutation(n), :]
    else:
        c_new = np.where(rng.random(c_old.shape) < 0.3, -c_old, c_old).astype(np.int8)
    c_new = np.asfortranarray(c_new, dtype=np.int8)
    _, _, eT2m, _ = o.get_hopping_exp()
    a = field.alpha

    def weight(c):
        w = 1.0
        for b in range(field.flavors):
            sgn = 1.0 if b == 0 else -1.0
            P = np.eye(n)
            for l in range(M):
                P = eT2m @ np.diag(np.exp(sgn * a * c[:, l])) @ P
            w *= np.linalg.det(np.eye(n) + P)
        return w

    ratio = weight(c_new) / weight(c_old)
    if field.flavors == 1:
        ratio = ratio * ratio * np.exp(-(a * c_new.sum() - a * c_old.sum()))
    expected = abs(ratio)
    o.set_uniforms(np.array([0.0]))            # accept whenever p <= 1 too
    acc, p = o.global_update(c_new)
    assert acc
    assert abs(p - expected) <= 1e-10 * max(expected, 1.0), (p, expected)
    # accepted: the field is the proposal and G is the stack rebuilt from it
    assert np.array_equal(o.get_field(), c_new)
    fresh, *_ = make_chain(oracle.OracleChain, pkg, 2, U_ref, 1.0, seed=31)
    fresh.set_hopping(model.hopping_matrix())
    fresh.set_field(c_new)
    fresh.build_stack()
    assert np.abs(o.greens() - fresh.greens()).max() < 1e-13
    # rejected (uniform 1): the old field is restored and G is untouched
    G_before = o.greens()
    o.set_uniforms(np.array([1.0]))
    acc2, p2 = o.global_update(c_old)
    if p2 <= 1.0:
        assert not acc2
        assert np.array_equal(o.get_field(), c_new)
        assert np.array_equal(o.greens(), G_before)
