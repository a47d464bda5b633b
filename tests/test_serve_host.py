"""Host-side behaviour of the serving loop (no GPU needed): argument checks, refusal of a
request larger than the memory budget, and a request that fails (no device here, or a device
the chain cannot be created on) reported as failed instead of hanging the loop."""
import numpy as np
import pytest

from conftest import make_chain


def test_server_create_rejects_bad_config(pkg, hip):
    with pytest.raises(hip.DQMCError):
        hip.Server(memory_budget=0, max_active=1)
    with pytest.raises(hip.DQMCError):
        hip.Server(memory_budget=1 << 30, max_active=0)


def test_server_refuses_request_over_budget_and_reports_failures(pkg, hip):
    g, model, par, field = make_chain(lambda *a, **k: None, pkg, 4, 4.0, 1.0, seed=3)
    fp = hip.footprint_bytes(len(model.l), field.flavors, par.slices, par.safe_mult)
    srv = hip.Server(memory_budget=fp, max_active=2, device=1 << 20)   # no such device
    try:
        T = model.hopping_matrix()
        with pytest.raises(hip.DQMCError):       # bad leading dimension / shapes
            srv.submit(T[:2, :2], field.conf, field.flavors, par.safe_mult, par.delta_tau,
                       field.alpha, 1, 1)
        g2, m2, p2, f2 = make_chain(lambda *a, **k: None, pkg, 6, 4.0, 1.0, seed=3)
        with pytest.raises(hip.DQMCError):       # footprint beyond the whole budget
            srv.submit(m2.hopping_matrix(), f2.conf, f2.flavors, p2.safe_mult, p2.delta_tau,
                       f2.alpha, 1, 1)
        ids = [srv.submit(T, field.conf, field.flavors, par.safe_mult, par.delta_tau,
                          field.alpha, 7 + k, 1) for k in range(3)]
        for rid in ids:
            srv.wait(rid)
            assert srv.poll(rid) == "failed"
            with pytest.raises(hip.DQMCError):
                srv.result(rid)
        st = srv.status()
        assert st["active"] == 0 and st["queued"] == 0 and st["used_bytes"] == 0
        assert st["peak_active"] == 1                # one footprint fits the budget
    finally:
        srv.close()
