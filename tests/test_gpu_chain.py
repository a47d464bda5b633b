"""Chain-level parity of the GPU sweep against the CPU oracle (same inputs, same uniforms).

Reference path: reverse_build_stack/propagate (stack.jl:284-308, 605-730), calculate_greens
(stack.jl:525-583), sweep_spatial (local_updates.jl:23-60).
Tolerances: G within 1e-10 relative (max-norm) in FP64; the accept sequence is bit-exact up
to the first proposal whose |u - p| lies within the tie band (1e-12), and the field and G
after the sweep are compared in full when there is no tie.
"""
import numpy as np
import pytest

from conftest import assert_same_decisions, counter_stream, make_chain, rel_maxdiff

pytestmark = pytest.mark.gpu

CASES = [
    # (L, U_ref, beta, mu): U_ref > 0 attractive (density field, 1 flavour), < 0 repulsive
    (4, 4.0, 4.0, 0.0),     # attractive 4x4, beta=4 (M=40): the CPU-path configuration
    (4, -4.0, 2.0, 0.0),    # repulsive 4x4 (magnetic, 2 flavours)
    (6, 4.0, 3.0, -0.5),    # attractive 6x6 with mu
    (3, -2.0, 1.5, 0.0),    # odd lattice, M=15 (uneven chunks)
]


def pair(pkg, hip, oracle, L, U, beta, mu, seed=1234, **kw):
    g, model, par, field = make_chain(hip.HIPChain, pkg, L, U, beta, mu=mu, seed=seed, **kw)
    o, *_ = make_chain(oracle.OracleChain, pkg, L, U, beta, mu=mu, seed=seed)
    T = model.hopping_matrix()
    o.set_hopping(T)
    g.set_hopping_exp(*o.get_hopping_exp())
    g.set_field(field.conf)
    o.set_field(field.conf)
    return g, o, model, par, field


@pytest.mark.parametrize("L,U,beta,mu", CASES)
def test_build_stack_greens(pkg, hip, oracle, L, U, beta, mu):
    g, o, *_ = pair(pkg, hip, oracle, L, U, beta, mu)
    g.build_stack()
    o.build_stack()
    assert g.state() == o.state()
    assert rel_maxdiff(g.greens(), o.greens()) < 1e-10
    assert rel_maxdiff(g.greens(physical=True), o.greens(physical=True)) < 1e-10


@pytest.mark.parametrize("L,U,beta,mu", CASES[:2])
def test_forward_build_equals_reverse_build(pkg, hip, oracle, L, U, beta, mu):
    """build_stack + M propagates == reverse_build_stack + propagate (flavortests_DQMC.jl:313-342)."""
    g, o, model, par, field = pair(pkg, hip, oracle, L, U, beta, mu)
    g.build_stack_forward()
    for _ in range(par.slices):
        g.propagate()
    g2, *_ = pair(pkg, hip, oracle, L, U, beta, mu)
    g2.build_stack()
    assert g.state() == g2.state()
    assert rel_maxdiff(g.greens(), g2.greens()) < 1e-10


@pytest.mark.parametrize("L,U,beta,mu", CASES)
def test_calculate_greens_from_scratch(pkg, hip, oracle, L, U, beta, mu):
    g, o, model, par, field = pair(pkg, hip, oracle, L, U, beta, mu)
    for sl in sorted({0, 1, par.slices // 3, par.slices - 1, par.slices}):
        assert rel_maxdiff(g.calculate_greens(sl), o.calculate_greens(sl)) < 1e-10, sl


def test_hopping_exponentials_on_device(pkg, hip, oracle):
    """exp(-+dtau T/2), exp(-+dtau T) computed on the GPU == eigen-decomposition oracle."""
    for lat in (pkg.SquareLattice(4), pkg.Honeycomb(3), pkg.SquareLattice(2)):
        model = pkg.HubbardModel(l=lat, U=1.0, mu=0.3)
        n = len(lat)
        g = hip.HIPChain(n, 1, 10, 10, 0.1, 0.5)
        o = oracle.OracleChain(n, 1, 10, 10, 0.1, 0.5)
        T = model.hopping_matrix()
        g.set_hopping(T)
        o.set_hopping(T)
        for a, b in zip(g.get_hopping_exp(), o.get_hopping_exp()):
            assert np.abs(a - b).max() < 1e-14
            assert np.array_equal(a, a.T)


@pytest.mark.parametrize("L,U,beta,mu", CASES)
@pytest.mark.parametrize("stream", ["supplied", "counter"])
def test_one_sweep_accept_sequence(pkg, hip, oracle, L, U, beta, mu, stream):
    g, o, model, par, field = pair(pkg, hip, oracle, L, U, beta, mu)
    n, M = len(model.l), par.slices
    cap = 2 * n * M
    if stream == "supplied":
        u = np.random.default_rng(99).random(cap)
        g.set_uniforms(u)
        o.set_uniforms(u)
    else:
        u = counter_stream(2024, cap)
        g.set_seed(2024)
        o.set_seed(2024)
    g.set_log(cap)
    o.set_log(cap)
    g.build_stack()
    o.build_stack()
    ag = g.sweep(1)
    ao = o.sweep(1)
    dg, pg = g.get_log()
    do, po = o.get_log()
    cg, co = g.counters(), o.counters()
    assert len(dg) == len(do) == cap
    if assert_same_decisions(dg, do, po, u):
        assert np.array_equal(dg, do)
        assert ag == ao
        assert np.array_equal(g.get_field(), o.get_field())
        assert cg["uniforms_used"] == co["uniforms_used"]
        assert rel_maxdiff(g.greens(), o.greens()) < 1e-10
    # acceptance probabilities agree to rounding
    assert np.abs(pg - po).max() < 1e-8 * max(1.0, np.abs(po).max())
    assert g.state() == o.state() == (1, 1, 1)


@pytest.mark.parametrize("L,U,beta,mu,lattice,window", [
    (10, -4.0, 1.0, 0.0, None, 0),      # n = 100: windows 32+32+32+4
    (10, 4.0, 1.0, -0.3, None, 16),     # 16-site windows
    (10, -3.0, 1.0, 0.0, None, 64),     # 64-site windows: 64 + 36
    (4, 3.0, 1.0, 0.1, "honeycomb", 0),  # honeycomb 2 L^2 = 32 sites
])
def test_window_widths_and_lattices(pkg, hip, oracle, L, U, beta, mu, lattice, window):
    """One sweep with supplied uniforms: decisions, field and G against the oracle for window
    widths 16 / 32 / 64, a lattice whose size is no multiple of the window, and honeycomb."""
    lat = pkg.Honeycomb(L) if lattice == "honeycomb" else None
    g, model, par, field = make_chain(hip.HIPChain, pkg, L, U, beta, mu=mu, lattice=lat, seed=21,
                                      delay_depth=window)
    o, *_ = make_chain(oracle.OracleChain, pkg, L, U, beta, mu=mu, lattice=lat, seed=21)
    o.set_hopping(model.hopping_matrix())
    g.set_hopping_exp(*o.get_hopping_exp())
    g.set_field(field.conf)
    o.set_field(field.conf)
    n, M = len(model.l), par.slices
    u = np.random.default_rng(7).random(2 * n * M)
    g.set_uniforms(u)
    o.set_uniforms(u)
    g.set_log(2 * n * M)
    o.set_log(2 * n * M)
    g.build_stack()
    o.build_stack()
    g.sweep(1)
    o.sweep(1)
    dg, _ = g.get_log()
    do, po = o.get_log()
    if assert_same_decisions(dg, do, po, u):
        assert np.array_equal(g.get_field(), o.get_field())
        assert rel_maxdiff(g.greens(), o.greens()) < 1e-10


def test_u0_closed_form(pkg, hip):
    """U = 0: G = [I + exp(-beta T)]^-1 at 1e-12 (flavortests_DQMC.jl:355-386)."""
    for model in (pkg.HubbardModel(7, 2, U=0.0, t=1.0), pkg.HubbardModel(8, 2, U=0.0, mu=1.0, t=1.0)):
        for beta in (1.0, 10.0):
            par = pkg.DQMCParameters.create(beta=beta, delta_tau=0.1, safe_mult=5)
            n = len(model.l)
            g = hip.HIPChain(n, 1, par.slices, 5, 0.1, 0.0)
            g.set_hopping(model.hopping_matrix())
            g.set_field(np.ones((n, par.slices), dtype=np.int8, order="F"))
            g.set_seed(1)
            g.build_stack()
            g.sweep(2)
            w, V = np.linalg.eigh(model.hopping_matrix())
            G = V @ np.diag(1.0 / (1.0 + np.exp(-beta * w))) @ V.T
            Gd = g.greens(physical=True)[0]
            assert np.abs(Gd - G).max() < 1e-12 + 1e-12 * np.abs(G).max()


def test_config_b_sweep_parity(pkg, hip, oracle):
    """Repulsive 16x16, beta=10 (N=256, M=100, 2 flavours): one full sweep against the oracle."""
    g, o, model, par, field = pair(pkg, hip, oracle, 16, -4.0, 10.0, 0.0, seed=77)
    n, M = 256, 100
    u = np.random.default_rng(5).random(2 * n * M)
    g.set_uniforms(u)
    o.set_uniforms(u)
    g.set_log(2 * n * M)
    o.set_log(2 * n * M)
    g.build_stack()
    o.build_stack()
    assert rel_maxdiff(g.greens(), o.greens()) < 1e-10
    g.sweep(1)
    o.sweep(1)
    dg, pg = g.get_log()
    do, po = o.get_log()
    mism = int((dg != do).sum())
    first = int(np.argmax(dg != do)) if mism else -1
    cg, co = g.counters(), o.counters()
    print("config B: mismatches", mism, "first", first, "ties", cg["ties"], co["ties"],
          "acc", cg["accepted"], co["accepted"], "prop_err", cg["prop_err_last"], co["prop_err_last"])
    # the device's wrap-vs-recompute check (max |G_wrapped - G_fresh| at the last
    # stabilisation) measures the same drift as the oracle's, to within rounding order
    assert 0.0 < cg["prop_err_last"] < 1e-6 and 0.0 < co["prop_err_last"] < 1e-6
    assert 0.1 < cg["prop_err_last"] / co["prop_err_last"] < 10.0
    if assert_same_decisions(dg, do, po, u):
        assert mism == 0
        assert np.array_equal(g.get_field(), o.get_field())
        assert rel_maxdiff(g.greens(), o.greens()) < 1e-10


@pytest.mark.parametrize("case", ["repulsive", "attractive_mu1"])
def test_dqmc_driver_on_gpu_matches_ed(pkg, case):
    """The reference's finite-U ED test (test/ED/ED_tests.jl:402-484) through the host driver
    on the HIP chain: 2x2, beta=1, dtau=0.1, safe_mult=5, 5k + 5k sweeps; atol 3.05 dtau^2,
    rtol 2 dtau^2."""
    import json
    import os
    from conftest import ROOT
    ed = json.load(open(os.path.join(ROOT, "tests", "golden", "ed_2x2.json")))[case]
    model = pkg.HubbardModel(2, 2, U=ed["U"], mu=ed["mu"], t=1.0)
    mc = pkg.DQMC(model, beta=1.0, delta_tau=0.1, safe_mult=5, thermalization=5000, sweeps=5000,
                  measure_rate=1, seed=77)
    assert mc.run(verbose=False) == pkg.ExitCode.SUCCESS
    ms = mc.measurements["equal_time"]
    atol, rtol = 3.05 * 0.01, 2 * 0.01
    close = lambda a, b: np.all(np.abs(np.asarray(a) - np.asarray(b)) <= atol + rtol * np.abs(np.asarray(b)))
    assert close(ms.mean("total_energy"), ed["energy"])
    G = ms.mean("greens")
    Ged = np.asarray(ed["greens"])
    for b in range(G.shape[0]):
        assert close(G[b], Ged[b * 4:(b + 1) * 4, b * 4:(b + 1) * 4])
    assert close(ms.mean("occupation"), 1 - np.diag(Ged)[: 4 * G.shape[0]])
    assert close(ms.mean("double_occupancy"), ed["double_occupancy"])


def test_batch_sweep_matches_chains_alone(pkg, hip):
    """Chains of different size swept together (dqmc_batch_sweep) give bit for bit what each
    gives alone: decisions, field and Green's function."""
    cases = [(4, 4.0, 2.0, 0.0, 11), (6, -4.0, 1.5, 0.0, 12), (8, 2.0, 1.0, -0.3, 13)]
    def build(L, U, beta, mu, seed):
        g, model, par, field = make_chain(hip.HIPChain, pkg, L, U, beta, mu=mu, seed=seed)
        g.set_hopping(model.hopping_matrix())
        g.set_field(field.conf)
        g.set_seed(seed)
        g.build_stack()
        return g
    alone = []
    for c in cases:
        g = build(*c)
        acc = g.sweep(3)
        alone.append((acc, g.get_field(), g.greens()))
        g.close()
    chains = [build(*c) for c in cases]
    accs = hip.batch_sweep(chains, 3)
    for (acc, f, G), g, a in zip(alone, chains, accs):
        assert a == acc
        assert np.array_equal(g.get_field(), f)
        assert np.array_equal(g.greens(), G)
        g.close()


@pytest.mark.parametrize("L,U,beta,mu", [(4, 4.0, 2.0, 0.0), (4, -4.0, 2.0, 0.0), (6, 2.0, 1.5, -0.4)])
def test_device_measurements_match_host(pkg, hip, L, U, beta, mu):
    """dqmc_measure_equal_time against the host kernels (measurements.py) on the same physical G."""
    from montecarlo_jl_amd import measurements as ms
    g, model, par, field = make_chain(hip.HIPChain, pkg, L, U, beta, mu=mu, seed=5)
    T = model.hopping_matrix()
    g.set_hopping(T)
    g.set_field(field.conf)
    g.set_seed(3)
    g.build_stack()
    g.sweep(1)
    G = g.greens(physical=True)
    d = g.measure_equal_time(U)
    tol = lambda a, b: abs(a - b) <= 1e-12 * max(1.0, abs(b))
    assert tol(d["density"], ms.density(G))
    assert tol(d["kinetic_energy"], ms.kinetic_energy(G, T))
    assert tol(d["interaction_energy"], ms.interaction_energy(G, U))
    assert tol(d["total_energy"], ms.total_energy(G, T, U))
    assert np.allclose(d["occupation"], ms.occupation(G), rtol=0, atol=1e-13)
    assert np.allclose(d["double_occupancy"], ms.double_occupancy(G), rtol=0, atol=1e-13)
    g.close()


def test_driver_device_measurements_match_host_path(pkg):
    """The run! loop with observables measured on the device equals the host path sample for
    sample (same seed, same chain)."""
    model = pkg.HubbardModel(4, 2, U=-4.0, mu=0.0, t=1.0)
    kw = dict(beta=2.0, delta_tau=0.1, safe_mult=10, thermalization=5, sweeps=20, measure_rate=2, seed=9)
    a = pkg.DQMC(model, **kw)
    b = pkg.DQMC(model, device_measurements=True, **kw)
    assert a.run(verbose=False) == pkg.ExitCode.SUCCESS
    assert b.run(verbose=False) == pkg.ExitCode.SUCCESS
    ma, mb = a.measurements["equal_time"], b.measurements["equal_time"]
    assert ma.count == mb.count > 0
    for k in ("density", "kinetic_energy", "interaction_energy", "total_energy"):
        assert abs(ma.mean(k) - mb.mean(k)) <= 1e-11 * max(1.0, abs(ma.mean(k)))
    assert np.allclose(ma.mean("double_occupancy"), mb.mean("double_occupancy"), rtol=0, atol=1e-12)


@pytest.mark.parametrize("L,beta", [(6, 1.5), (16, 10.0), (3, 0.7)])
def test_field_packed_on_device(pkg, hip, L, beta):
    """The field as BitArray chunks (compress / decompress!, fields.jl:331-334), packed and
    unpacked on the device, exact and ragged sizes; a chain resumed from the chunks builds
    the same G."""
    g, model, par, field = make_chain(hip.HIPChain, pkg, L, -4.0, beta, seed=41)
    h = None
    try:
        g.set_hopping(model.hopping_matrix())
        g.set_field(field.conf)
        assert np.array_equal(g.get_field_packed(), field.compress_chunks())
        g.set_seed(41)
        g.build_stack()
        g.sweep(1)
        chunks = g.get_field_packed()
        assert np.array_equal(chunks, pkg.fields.bits_to_chunks(g.get_field() == 1))
        # resume in a fresh chain from the chunks
        h, _, _, _ = make_chain(hip.HIPChain, pkg, L, -4.0, beta, seed=41)
        h.set_hopping(model.hopping_matrix())
        h.set_field_packed(chunks)
        assert np.array_equal(h.get_field(), g.get_field())
        g.build_stack()
        h.build_stack()
        assert np.array_equal(h.greens(), g.greens())
        # refusals: wrong chunk count, bits beyond the field in the last chunk
        with pytest.raises(hip.DQMCError):
            h.set_field_packed(chunks[:-1])
        count = g.n * g.M
        if count % 64:
            bad = chunks.copy()
            bad[-1] |= np.uint64(1) << np.uint64(63)
            with pytest.raises(hip.DQMCError):
                h.set_field_packed(bad)
    finally:
        g.close()
        if h is not None:
            h.close()
