"""Global updates on the device (dqmc_global_update) against the oracle's restatement of
global_update (global_updates.jl:136-207), itself pinned by brute force in test_oracle.py."""
import numpy as np
import pytest

from conftest import make_chain, rel_maxdiff

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("L,U,beta,mu", [(4, 4.0, 2.0, 0.0), (4, -4.0, 2.0, 0.0), (6, 3.0, 1.5, -0.3),
                                         (8, -3.0, 3.0, 0.0)])
def test_global_moves_match_oracle(pkg, hip, oracle, L, U, beta, mu):
    g, model, par, field = make_chain(hip.HIPChain, pkg, L, U, beta, mu=mu, seed=9)
    o, *_ = make_chain(oracle.OracleChain, pkg, L, U, beta, mu=mu, seed=9)
    try:
        o.set_hopping(model.hopping_matrix())
        g.set_hopping_exp(*o.get_hopping_exp())
        for c in (g, o):
            c.set_field(field.conf)
            c.set_seed(77)
            c.build_stack()
        rng = np.random.default_rng(3)
        n, M = len(model.l), par.slices
        moves = []
        tied = False
        for k in range(6):
            c = g.get_field()
            kind = k % 3
            if kind == 0:
                prop = -c
            elif kind == 1:
                prop = c[rng.permutation(n), :]
            else:
                prop = np.where(rng.random(c.shape) < 0.02, -c, c)
            prop = np.asfortranarray(prop, dtype=np.int8)
            ag, pg = g.global_update(prop)
            ao, po = o.global_update(prop)
            moves.append((ag, pg, ao, po))
            assert abs(pg - po) <= 1e-9 * max(po, 1e-300) + 1e-300, (k, pg, po)
            if abs(po - 1.0) < 1e-8:
                # a symmetry of the model makes the weight ratio 1 up to rounding (e.g. the
                # flip of a magnetic field swaps the spin blocks): whether a uniform is drawn
                # (p <= 1) is a tie, so the two chains may part here
                tied = True
                break
            assert ag == ao
            assert np.array_equal(g.get_field(), o.get_field())
            assert g.state() == o.state() == (1, 1, 1)
            assert rel_maxdiff(g.greens(), o.greens()) < 1e-10
            # a sweep in between keeps the chains together and brings the state back to (1, 1, 1)
            g.sweep(1)
            o.sweep(1)
        print(moves)
        if not tied:
            assert g.counters()["uniforms_used"] == o.counters()["uniforms_used"]
    finally:
        g.close()
        o.close()


def test_global_update_refuses_other_states(pkg, hip):
    g, model, par, field = make_chain(hip.HIPChain, pkg, 4, 4.0, 2.0, seed=9)
    try:
        g.set_hopping(model.hopping_matrix())
        g.set_field(field.conf)
        g.build_stack()
        g.sweep_spatial()
        g.propagate()                       # current_slice 2
        with pytest.raises(hip.DQMCError, match="current_slice"):
            g.global_update(-field.conf)
        g.build_stack()
        g.calculate_greens(3)               # the stack's D no longer belongs to G
        with pytest.raises(hip.DQMCError, match="current_slice"):
            g.global_update(-field.conf)
        g.build_stack()
        bad = field.conf.copy()
        bad[0, 0] = 0
        with pytest.raises(hip.DQMCError):
            g.global_update(bad)
    finally:
        g.close()
