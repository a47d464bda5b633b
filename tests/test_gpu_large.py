"""Chains beyond N = 256 on the GPU: config C (attractive 24x24, N = 576, mu = -0.5,
beta = 12, M = 120) against the oracle, and the serving mix at config E's largest shapes
(N = 576, 1024 and honeycomb L = 32, N = 2048, two flavours), each request equal to its
chain run alone."""
import time

import numpy as np
import pytest

from conftest import assert_same_decisions, make_chain, rel_maxdiff

pytestmark = pytest.mark.gpu


def test_config_c_fp64_against_oracle(pkg, hip, oracle):
    """Stack build (G at slice 1) and the first 12 slice steps (sweep_spatial + propagate,
    through the stabilisation at slice 10) with supplied uniforms: G within 1e-10, decisions
    bit-exact up to the first tie."""
    t0 = time.time()
    g, model, par, field = make_chain(hip.HIPChain, pkg, 24, 4.0, 12.0, mu=-0.5, seed=3)
    o, *_ = make_chain(oracle.OracleChain, pkg, 24, 4.0, 12.0, mu=-0.5, seed=3)
    try:
        n, M = 576, 120
        assert par.slices == M
        o.set_hopping(model.hopping_matrix())
        g.set_hopping_exp(*o.get_hopping_exp())
        g.set_field(field.conf)
        o.set_field(field.conf)
        steps = 12
        u = np.random.default_rng(8).random(steps * n)
        for c in (g, o):
            c.set_uniforms(u)
            c.set_log(steps * n)
            c.build_stack()
        print(f"config C build: {time.time() - t0:.1f} s", flush=True)
        assert g.state() == o.state() == (1, 1, 1)
        assert rel_maxdiff(g.greens(), o.greens()) < 1e-10
        for _ in range(steps):
            for c in (g, o):
                c.sweep_spatial()
                c.propagate()
        print(f"config C steps: {time.time() - t0:.1f} s", flush=True)
        assert g.state() == o.state() == (steps + 1, 1, 2)
        dg, _ = g.get_log()
        do, po = o.get_log()
        if assert_same_decisions(dg, do, po, u):
            assert np.array_equal(g.get_field(), o.get_field())
            assert rel_maxdiff(g.greens(), o.greens()) < 1e-10
        cg, co = g.counters(), o.counters()
        print("config C: accepted", cg["accepted"], co["accepted"], "prop err", cg["prop_err_last"],
              co["prop_err_last"])
    finally:
        g.close()
        o.close()


CONFIG_E_SHAPES = [
    # (lattice, L, U_ref, beta, seed): config E's largest chains, beta cut to keep the test short
    ("square", 24, 3.0, 1.0, 101),      # N = 576, density field
    ("square", 32, -5.0, 0.6, 102),     # N = 1024, magnetic field (two flavours)
    ("honeycomb", 32, -2.0, 0.5, 103),  # N = 2048, magnetic field (two flavours)
]


def test_serving_large_shapes_equal_chains_alone(pkg, hip):
    """Requests of N = 576, 1024 and 2048 run together in the server give bit for bit what
    each chain gives alone (field, G, accepted flips) after a stack build and one sweep."""
    reqs = []
    for lat, L, U, beta, seed in CONFIG_E_SHAPES:
        l = pkg.Honeycomb(L) if lat == "honeycomb" else pkg.SquareLattice(L)
        model = pkg.HubbardModel(l=l, U=U)
        par = pkg.DQMCParameters.create(beta=beta, delta_tau=0.1, safe_mult=10)
        field = model.choose_field()(par, model)
        field.rand(seed)
        reqs.append((model, par, field, seed))
    alone = []
    for model, par, field, seed in reqs:
        t0 = time.time()
        n = len(model.l)
        g = hip.HIPChain(n, field.flavors, par.slices, 10, par.delta_tau, field.alpha)
        g.set_hopping(model.hopping_matrix())
        g.set_field(field.conf)
        g.set_seed(seed)
        g.build_stack()
        acc = g.sweep(1)
        alone.append((acc, g.get_field(), g.greens()))
        g.close()
        print(f"alone N={n} flavours={field.flavors} M={par.slices}: {time.time() - t0:.2f} s", flush=True)
    total = sum(hip.footprint_bytes(len(m.l), f.flavors, p.slices, 10) for m, p, f, _ in reqs)
    srv = hip.Server(memory_budget=total, max_active=3)
    try:
        ids = [srv.submit(m.hopping_matrix(), f.conf, f.flavors, 10, p.delta_tau, f.alpha, seed, 1)
               for m, p, f, seed in reqs]
        for rid in ids:
            srv.wait(rid)
        assert srv.status()["peak_active"] >= 2
        for rid, (acc, conf, G) in zip(ids, alone):
            r = srv.result(rid)
            assert r["accepted"] == acc
            assert np.array_equal(r["field"], conf)
            assert np.array_equal(r["greens"], G)
    finally:
        srv.close()
