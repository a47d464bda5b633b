"""Serving form (dqmc_server_*): requests of uneven shape admitted by HBM footprint, running
concurrently; each result must equal the same chain run alone, bit for bit."""
import time

import numpy as np
import pytest

from conftest import make_chain

pytestmark = pytest.mark.gpu

# (L, U_ref, beta, mu, lattice, seed, nsweeps): square and honeycomb, density and magnetic
CASES = [
    (4, 4.0, 2.0, 0.0, None, 21, 3),
    (6, -4.0, 1.5, 0.0, None, 22, 2),
    (8, 2.0, 1.0, -0.3, None, 23, 2),
    (3, 3.0, 1.0, 0.1, "honeycomb", 24, 3),
    (4, -2.0, 3.0, 0.2, None, 25, 2),
]


def _lattice(pkg, name, L):
    return None if name is None else pkg.Honeycomb(L)


def _alone(hip, pkg, case):
    L, U, beta, mu, lat, seed, ns = case
    g, model, par, field = make_chain(hip.HIPChain, pkg, L, U, beta, mu=mu,
                                      lattice=_lattice(pkg, lat, L), seed=seed)
    g.set_hopping(model.hopping_matrix())
    g.set_field(field.conf)
    g.set_seed(seed)
    g.build_stack()
    acc = g.sweep(ns)
    out = (acc, g.get_field(), g.greens())
    g.close()
    return out


def _submit(srv, hip, pkg, case):
    L, U, beta, mu, lat, seed, ns = case
    g, model, par, field = make_chain(lambda *a, **k: None, pkg, L, U, beta, mu=mu,
                                      lattice=_lattice(pkg, lat, L), seed=seed)
    return srv.submit(model.hopping_matrix(), field.conf, field.flavors, par.safe_mult,
                      par.delta_tau, field.alpha, seed, ns)


def test_server_results_equal_chains_alone(pkg, hip):
    alone = [_alone(hip, pkg, c) for c in CASES]
    srv = hip.Server(memory_budget=1 << 34, max_active=3)
    try:
        ids = [_submit(srv, hip, pkg, c) for c in CASES[:2]]
        time.sleep(0.05)                       # later requests arrive while these run
        ids += [_submit(srv, hip, pkg, c) for c in CASES[2:]]
        for rid in ids:
            srv.wait(rid)
            assert srv.poll(rid) == "done"
        st = srv.status()
        assert 1 <= st["peak_active"] <= 3
        for rid, (acc, f, G) in zip(ids, alone):
            r = srv.result(rid)
            assert r["accepted"] == acc
            assert np.array_equal(r["field"], f)
            assert np.array_equal(r["greens"], G)
            assert r["latency_s"] >= r["queued_s"] >= 0.0
        assert srv.status()["active"] == 0 and srv.status()["used_bytes"] == 0
    finally:
        srv.close()


def test_server_admission_by_footprint(pkg, hip):
    case = (6, 4.0, 1.0, 0.0, None, 31, 1)
    g, model, par, field = make_chain(lambda *a, **k: None, pkg, 6, 4.0, 1.0, seed=31)
    fp = hip.footprint_bytes(len(model.l), field.flavors, par.slices, par.safe_mult)
    srv = hip.Server(memory_budget=int(2.5 * fp), max_active=8)
    try:
        ids = [_submit(srv, hip, pkg, case) for _ in range(5)]
        for rid in ids:
            srv.wait(rid)
        st = srv.status()
        assert st["peak_active"] <= 2            # two footprints fit the budget, three do not
        ref = _alone(hip, pkg, case)
        for rid in ids:
            r = srv.result(rid)
            assert r["accepted"] == ref[0]
            assert np.array_equal(r["field"], ref[1])
        # a request larger than the whole budget is refused at submission
        big = (16, 4.0, 2.0, 0.0, None, 32, 1)
        with pytest.raises(hip.DQMCError):
            _submit(srv, hip, pkg, big)
    finally:
        srv.close()
