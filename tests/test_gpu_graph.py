"""A sweep replayed from its captured HIP graph equals the sweep launched kernel by kernel,
bit for bit, replay after replay, with eager work, state changes and synchronisations in
between (which force recaptures where the baked-in host state changed)."""
import numpy as np
import pytest

from conftest import make_chain

pytestmark = pytest.mark.gpu


def _pair(pkg, hip, L, U, beta, seed=21):
    chains = []
    for graph in (True, False):
        g, model, par, field = make_chain(hip.HIPChain, pkg, L, U, beta, seed=seed)
        g.set_hopping(model.hopping_matrix())
        g.set_field(field.conf)
        g.set_seed(99)
        g.set_sweep_graph(graph)
        g.set_log(8 * 2 * len(model.l) * par.slices)
        g.build_stack()
        chains.append(g)
    return chains, model, par


@pytest.mark.parametrize("L,U,beta", [(4, 4.0, 2.0), (6, -4.0, 2.0), (12, -4.0, 1.0)])
def test_graph_replays_equal_eager_sweeps(pkg, hip, L, U, beta):
    (gg, ge), model, par = _pair(pkg, hip, L, U, beta)
    try:
        for it in range(5):
            if it == 2:
                # eager work in between, then a multi-sweep call
                for c in (gg, ge):
                    c.calculate_greens(3)
                    c.greens(physical=True)
                    c.synchronize()
                ag, ae = gg.sweep(2), ge.sweep(2)
            elif it == 3:
                # a full cycle of propagates moves G; the state returns to the start
                for c in (gg, ge):
                    for _ in range(2 * par.slices):
                        c.propagate()
                ag, ae = gg.sweep(1), ge.sweep(1)
            elif it == 4:
                # a new seed changes a baked-in argument: recapture
                for c in (gg, ge):
                    c.set_seed(7, 0)
                ag, ae = gg.sweep(1), ge.sweep(1)
            else:
                ag, ae = gg.sweep(1), ge.sweep(1)
            assert ag == ae, it
            assert gg.state() == ge.state(), it
            assert np.array_equal(gg.get_field(), ge.get_field()), it
            assert np.array_equal(gg.greens(), ge.greens()), it
            dg, pg = gg.get_log()
            de, pe = ge.get_log()
            assert np.array_equal(dg, de) and np.array_equal(pg, pe), it
        sg, se = gg.counters(), ge.counters()
        assert sg == se
        assert np.array_equal(gg.stabilization_log(), ge.stabilization_log())
    finally:
        gg.close()
        ge.close()
