"""The many-chain form: dqmc_batch_sweep runs FP64 chains in lockstep slice visits, one
launch per kernel kind covering every chain.  Each chain's result must equal the chain swept
alone bit for bit (decisions, field, Green's function, statistics)."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def config_d_chains(pkg, count, seed=2024, Ls=(6, 8, 10, 12, 16), betas=(4.0, 6.0, 8.0, 10.0)):
    """Config D's distribution: L uniform in {6, 8, 10, 12, 16} (square), U uniform in [-8, 8]
    (this package's reference convention: negative = repulsive, magnetic field, 2 flavours),
    beta in {4, 6, 8, 10}, dtau = 0.1, mu = 0, a seeded +-1 field per chain."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        L = int(rng.choice(Ls))
        U = float(rng.uniform(-8.0, 8.0)) or 1.0
        beta = float(rng.choice(betas))
        model = pkg.HubbardModel(L, 2, U=U)
        par = pkg.DQMCParameters.create(beta=beta, delta_tau=0.1, safe_mult=10)
        field = model.choose_field()(par, model)
        field.rand(100 + k)
        out.append((model, par, field, 500 + k))
    return out


def make(hip, model, par, field, seed):
    n = len(model.l)
    g = hip.HIPChain(n, field.flavors, par.slices, par.safe_mult, par.delta_tau, field.alpha)
    g.set_hopping(model.hopping_matrix())
    g.set_field(field.conf)
    g.set_seed(seed)
    g.build_stack()
    return g


def test_config_d_batch_equals_chains_alone(pkg, hip):
    specs = config_d_chains(pkg, 64)
    nsweeps = 2
    t0 = time.time()
    alone = []
    for model, par, field, seed in specs:
        g = make(hip, model, par, field, seed)
        acc = g.sweep(nsweeps)
        st = g.stats()
        alone.append((acc, g.get_field(), g.greens(), st, g.state()))
        g.close()
    t_alone = time.time() - t0
    chains = [make(hip, *s) for s in specs]
    try:
        t0 = time.time()
        accs = hip.batch_sweep(chains, nsweeps)
        t_batch = time.time() - t0
        print(f"64 chains x {nsweeps} sweeps: alone {t_alone:.2f} s (incl. builds), batched {t_batch:.2f} s")
        for (acc, f, G, st, state), g, a in zip(alone, chains, accs):
            assert a == acc
            assert g.state() == state
            assert np.array_equal(g.get_field(), f)
            assert np.array_equal(g.greens(), G)
            s2 = g.stats()
            for key in ("accepted", "proposed", "uniforms_used", "ties", "stabilizations",
                        "propagation_error_count", "sweeps", "device_errors"):
                assert s2[key] == st[key], key
            assert s2["propagation_error_max"] == st["propagation_error_max"]
        # a second batched call continues every chain where it stopped
        accs2 = hip.batch_sweep(chains, 1)
        assert all(a >= 0 for a in accs2)
    finally:
        for g in chains:
            g.close()


def test_batch_with_uneven_chunks_and_windows(pkg, hip):
    """Slice counts that are no multiple of safe_mult (uneven chunks, stabilisations at
    different steps), window widths 16 / 32 / 64 and honeycomb chains in one batch."""
    specs = []
    for k, (lat, L, U, beta, dd) in enumerate([("square", 5, 3.0, 1.3, 0), ("square", 7, -4.0, 2.1, 16),
                                                ("honeycomb", 4, -2.0, 1.7, 0), ("square", 10, 5.0, 0.9, 64),
                                                ("square", 3, -1.0, 0.7, 0)]):
        l = pkg.Honeycomb(L) if lat == "honeycomb" else pkg.SquareLattice(L)
        model = pkg.HubbardModel(l=l, U=U)
        par = pkg.DQMCParameters.create(beta=beta, delta_tau=0.1, safe_mult=4, warn_round=False)
        field = model.choose_field()(par, model)
        field.rand(7 + k)
        specs.append((model, par, field, 70 + k, dd))

    def mk(model, par, field, seed, dd):
        n = len(model.l)
        g = hip.HIPChain(n, field.flavors, par.slices, par.safe_mult, par.delta_tau, field.alpha,
                         delay_depth=dd)
        g.set_hopping(model.hopping_matrix())
        g.set_field(field.conf)
        g.set_seed(seed)
        g.build_stack()
        return g

    alone = []
    for s in specs:
        g = mk(*s)
        acc = g.sweep(3)
        alone.append((acc, g.get_field(), g.greens()))
        g.close()
    chains = [mk(*s) for s in specs]
    try:
        accs = hip.batch_sweep(chains, 3)
        for (acc, f, G), g, a in zip(alone, chains, accs):
            assert a == acc
            assert np.array_equal(g.get_field(), f)
            assert np.array_equal(g.greens(), G)
    finally:
        for g in chains:
            g.close()
