"""Time-displaced Green's functions on the device (dqmc_unequal_time_greens) against the
oracle's restatement of the unequal-time stack (oracle/unequal_time.py, pinned on the CPU by
test_unequal_time_oracle.py).  FP64 throughout: G within 1e-10 (relative to max |G|)."""
import numpy as np
import pytest

from conftest import rel_maxdiff

pytestmark = pytest.mark.gpu


def _pair(pkg, hip, oracle, L, dims, U, beta, safe_mult, seed=5, mu=0.0):
    model = pkg.HubbardModel(L, dims, U=U, mu=mu)
    par = pkg.DQMCParameters.create(beta=beta, delta_tau=0.1, safe_mult=safe_mult)
    field = model.choose_field()(par, model)
    field.rand(seed)
    n = len(model.l)
    o = oracle.OracleChain(n, field.flavors, par.slices, safe_mult, par.delta_tau, field.alpha)
    g = hip.HIPChain(n, field.flavors, par.slices, safe_mult, par.delta_tau, field.alpha)
    o.set_hopping(model.hopping_matrix())
    g.set_hopping_exp(*o.get_hopping_exp())
    o.set_field(field.conf)
    g.set_field(field.conf)
    return g, o, model, par, field


def _pairs(M, sm):
    ks = sorted({0, 1, sm - 1, sm, sm + 1, M // 2, M - 1, M} & set(range(M + 1)))
    out = [(k, l) for k in ks for l in ks]
    return out + [(7, 6), (6, 7)] if M > 7 else out


@pytest.mark.parametrize("L,dims,U,beta,sm", [
    (6, 1, 1.0, 5.0, 10),     # the reference's BigFloat case (unequal_time_stack.jl:174-300)
    (2, 2, 4.0, 2.0, 5),      # attractive: one flavour block
    (2, 2, -4.0, 3.0, 10),    # repulsive: two flavour blocks
    (4, 2, -2.0, 4.0, 8),
])
def test_matches_oracle(pkg, hip, oracle, L, dims, U, beta, sm):
    from oracle.unequal_time import UnequalTimeChain
    g, o, model, par, field = _pair(pkg, hip, oracle, L, dims, U, beta, sm)
    try:
        M = par.slices
        ut = UnequalTimeChain(o)
        for k, l in _pairs(M, sm):
            for phys in (False, True):
                Gg = g.unequal_time_greens(k, l, physical=phys)
                Go = ut.greens(k, l, physical=phys)
                assert rel_maxdiff(Gg, Go) < 1e-10, (k, l, phys, rel_maxdiff(Gg, Go))
    finally:
        g.close()
        o.close()


def test_identities_on_device(pkg, hip, oracle):
    """G(k, k) = calculate_greens(k) and G(k, 0) = -G(k, M) (unequal_time_stack.jl:88-102)."""
    g, o, model, par, field = _pair(pkg, hip, oracle, 6, 1, 1.0, 15.0, 5)
    try:
        M = par.slices
        for k in range(0, M + 1, 13):
            assert np.abs(g.unequal_time_greens(k, k, physical=False) - g.calculate_greens(k)).max() < 1e-13
        for k in range(0, M, 17):
            Gk0 = g.unequal_time_greens(k, 0)
            GkM = g.unequal_time_greens(k, M)
            assert np.abs(Gk0 + GkM).max() < 1e-13, k
    finally:
        g.close()
        o.close()


def test_follows_the_field_and_keeps_the_chain(pkg, hip, oracle):
    """The stack is rebuilt after a sweep changes the field; the chain's G, state and global
    updates are unaffected by the call."""
    from oracle.unequal_time import unequal_time_greens
    g, o, model, par, field = _pair(pkg, hip, oracle, 4, 2, -4.0, 2.0, 5)
    try:
        g.set_seed(3)
        g.build_stack()
        g.sweep(1)
        G0, st0 = g.greens(), g.state()
        Ga = g.unequal_time_greens(9, 4)
        assert np.array_equal(g.greens(), G0) and g.state() == st0
        o.set_field(g.get_field())
        assert rel_maxdiff(Ga, unequal_time_greens(o, 9, 4)) < 1e-10
        g.sweep(1)
        Gb = g.unequal_time_greens(9, 4)
        o.set_field(g.get_field())
        assert rel_maxdiff(Gb, unequal_time_greens(o, 9, 4)) < 1e-10
        assert rel_maxdiff(Ga, Gb) > 1e-6          # the field moved, so did G(9, 4)
        acc, p = g.global_update(np.asfortranarray(-g.get_field()))
        assert np.isfinite(p) and p > 0
    finally:
        g.close()
        o.close()


def test_bad_arguments(pkg, hip, oracle):
    g, o, model, par, field = _pair(pkg, hip, oracle, 2, 2, 4.0, 1.0, 5)
    try:
        for k, l in [(-1, 0), (0, par.slices + 1), (par.slices + 1, 3)]:
            with pytest.raises(hip.DQMCError):
                g.unequal_time_greens(k, l)
    finally:
        g.close()
        o.close()


def test_large_chain(pkg, hip, oracle):
    """n = 256 sites (16 x 16), two flavour blocks, beta 4."""
    from oracle.unequal_time import UnequalTimeChain
    g, o, model, par, field = _pair(pkg, hip, oracle, 16, 2, -4.0, 4.0, 10)
    try:
        ut = UnequalTimeChain(o)
        for k, l in [(23, 11), (11, 23), (40, 0), (0, 40)]:
            Gg = g.unequal_time_greens(k, l)
            Go = ut.greens(k, l)
            assert rel_maxdiff(Gg, Go) < 1e-10, (k, l, rel_maxdiff(Gg, Go))
    finally:
        g.close()
        o.close()


@pytest.mark.parametrize("L,dims,U,beta,sm,start,stop,rec", [
    (6, 1, 1.0, 15.0, 5, 0, 150, 5),      # the reference's iterator case (:141-170)
    (6, 1, 1.0, 15.0, 5, 0, 150, 20),
    (2, 2, -4.0, 4.0, 10, 1, 40, 10),
    (2, 2, -4.0, 4.0, 10, 7, 33, 20),
    (4, 2, 3.0, 3.0, 10, 0, 30, 20),
])
def test_combined_iterator_matches_oracle(pkg, hip, oracle, L, dims, U, beta, sm, start, stop, rec):
    """Between recalculations the iterator advances by slice products whose rounding grows
    with their condition (at U = -4 and recalculate = 2 safe_mult the oracle's own iterator
    departs from the direct G by up to ~1e-8, a bound the reference's verify() puts at 1e-6):
    the device must match the oracle within 1e-10 or within 4x the oracle iterator's own
    departure from the direct G at that slice, whichever is larger."""
    from oracle.unequal_time import UnequalTimeChain, combined_greens
    g, o, model, par, field = _pair(pkg, hip, oracle, L, dims, U, beta, sm)
    try:
        tg = g.combined_greens(start, stop, rec)
        to = combined_greens(o, start, stop, rec)
        ut = UnequalTimeChain(o)
        assert len(tg) == len(to) == stop - start + 1
        for i, (a, b) in enumerate(zip(tg, to)):
            l = start + i
            direct = (ut.greens(0, l), ut.greens(l, 0), ut.greens(l, l))
            drift = max(rel_maxdiff(b[j], direct[j]) for j in range(3))
            tol = max(1e-10, 4 * drift)
            for j in range(3):
                assert rel_maxdiff(a[j], b[j]) < tol, (l, j, rel_maxdiff(a[j], b[j]), tol)
    finally:
        g.close()
        o.close()


def test_combined_iterator_uses_the_chain_g(pkg, hip, oracle):
    """At current slice 1 the iterator starts from the chain's own G (greens_iterators.jl:
    197-200): after a sweep the chain is back there."""
    from oracle.unequal_time import combined_greens
    g, o, model, par, field = _pair(pkg, hip, oracle, 4, 2, -4.0, 2.0, 5)
    try:
        g.set_seed(3)
        g.build_stack()
        g.sweep(1)
        assert g.state()[0] == 1
        tg = g.combined_greens(0, par.slices, 10)
        o.set_field(g.get_field())
        to = combined_greens(o, 0, par.slices, 10, G00=g.greens())
        for i, (a, b) in enumerate(zip(tg, to)):
            for j in range(3):
                assert rel_maxdiff(a[j], b[j]) < 1e-10, (i, j)
        with pytest.raises(hip.DQMCError):
            g.combined_greens(5, 3, 10)
    finally:
        g.close()
        o.close()


@pytest.mark.parametrize("lat,U,beta,sm", [
    ("square4", -4.0, 2.0, 5),     # two flavour blocks (BlockDiagonal kernels)
    ("square4", 4.0, 3.0, 10),     # one block (DiagonallyRepeating kernels)
    ("honeycomb3", -2.0, 2.0, 10),  # two basis sites
])
def test_density_susceptibilities_match_oracle(pkg, hip, oracle, lat, U, beta, sm):
    from oracle.unequal_time import density_susceptibility
    l = pkg.SquareLattice(4, 4) if lat == "square4" else pkg.Honeycomb(3, 3)
    model = pkg.HubbardModel(l=l, U=U)
    par = pkg.DQMCParameters.create(beta=beta, delta_tau=0.1, safe_mult=sm)
    field = model.choose_field()(par, model)
    field.rand(8)
    n = len(l)
    o = oracle.OracleChain(n, field.flavors, par.slices, sm, par.delta_tau, field.alpha)
    g = hip.HIPChain(n, field.flavors, par.slices, sm, par.delta_tau, field.alpha)
    try:
        o.set_hopping(model.hopping_matrix())
        g.set_hopping_exp(*o.get_hopping_exp())
        o.set_field(field.conf)
        g.set_field(field.conf)
        g.set_seed(5)
        g.build_stack()
        g.sweep(1)
        o.set_field(g.get_field())
        Gcur = g.greens(physical=True)
        for obs in ("charge", "x", "y", "z"):
            for reduced in (False, True):
                for rec in (sm, 2 * sm):
                    dev = g.density_susceptibility(l, obs, recalculate=rec, reduced=reduced)
                    ref = density_susceptibility(o, l, Gcur, obs, recalculate=rec, reduced=reduced,
                                                 G00=g.greens())
                    assert dev.shape == ref.shape == (l.n_bravais, l.n_basis, l.n_basis)
                    if rec == sm:
                        tol, stable = 1e-10, ref
                    else:
                        # the iterator's own drift between recalculations (see above)
                        tol = max(1e-10, 4 * rel_maxdiff(ref, stable))
                    assert rel_maxdiff(dev, ref) < tol, (obs, reduced, rec, rel_maxdiff(dev, ref), tol)
    finally:
        g.close()
        o.close()


def test_density_susceptibilities_ed_u0(pkg, hip, oracle):
    """U = 0 on Honeycomb(2, 1) against exact diagonalisation (test/ED/ED_tests.jl:300-345)."""
    from oracle.ed import HubbardED
    from test_unequal_time_oracle import _ed_density
    l = pkg.Honeycomb(2, 1)
    model = pkg.HubbardModel(l=l, U=0.0)
    beta = 8.9
    par = pkg.DQMCParameters.create(beta=beta, delta_tau=0.1, safe_mult=5)
    field = pkg.MagneticHirschField(par, model)
    field.rand(1)
    n = len(l)
    g = hip.HIPChain(n, field.flavors, par.slices, 5, par.delta_tau, field.alpha)
    try:
        g.set_hopping(model.hopping_matrix())
        g.set_field(field.conf)
        g.build_stack()
        ed = HubbardED(model, beta)
        for obs in ("charge", "x", "y", "z"):
            dev = g.density_susceptibility(l, obs, recalculate=10)
            ref = _ed_density(ed, l, obs, par.delta_tau)
            assert np.allclose(dev, ref, atol=1e-12, rtol=1e-12), (obs, np.abs(dev - ref).max())
        with pytest.raises(hip.DQMCError):
            g.density_susceptibility(pkg.Honeycomb(3, 1), "charge")
    finally:
        g.close()


def test_driver_time_integral_measurements_u0(pkg, hip):
    """mc[:CDS] = charge_density_susceptibility(mc, model), mc[:SDSz] = ..., mc[:UTG] =
    greens_measurement(mc, model, GreensAt(3, 1)) through run!: at U = 0 every configuration
    gives the same values, so the means equal exact diagonalisation (ED_tests.jl:100-345)."""
    from oracle.ed import HubbardED
    from test_unequal_time_oracle import _ed_density
    ms = pkg.measurements
    l = pkg.Honeycomb(2, 1)
    model = pkg.HubbardModel(l=l, U=0.0)
    mc = pkg.DQMC(model, beta=2.0, delta_tau=0.1, safe_mult=5, thermalization=1, sweeps=4,
                  measure_rate=2, seed=3)
    mc["CDS"] = ms.charge_density_susceptibility(mc, model)
    mc["SDSz"] = ms.spin_density_susceptibility(mc, model, "z")
    mc["UTG"] = ms.greens_measurement(mc, model, ms.GreensAt(3, 1))
    mc.run(verbose=False)
    assert mc["CDS"].count == mc["SDSz"].count == mc["UTG"].count == 2
    ed = HubbardED(model, 2.0)
    for key, obs in (("CDS", "charge"), ("SDSz", "z")):
        ref = _ed_density(ed, l, obs, 0.1)
        assert np.allclose(mc[key].mean(), ref, atol=1e-12, rtol=1e-12), key
    assert mc["UTG"].mean().shape == (1, len(l), len(l))
