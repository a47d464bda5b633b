"""A slice's local updates run as one launch (the windows back to back, each window's update
of G applied beside the next window's decisions, ordered by device-side counters) and give the
same chain as one launch per window, bit for bit: the same decisions, field, G, counters and
stabilisation log, sweep after sweep, through graph replays and rebuilds."""
import numpy as np
import pytest

from conftest import make_chain

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("L,U,beta,W", [(4, 4.0, 2.0, 0), (6, 4.0, 2.0, 0), (12, 4.0, 1.0, 32),
                                        (16, 4.0, 1.0, 0), (8, -4.0, 2.0, 32), (6, 4.0, 2.0, 32),
                                        (16, 4.0, 1.0, 32), (24, -4.0, 1.0, 0), (24, -4.0, 1.0, 16)])
def test_slice_launch_equals_window_launches(pkg, hip, L, U, beta, W):
    chains = []
    try:
        for one in (True, False):
            g, model, par, field = make_chain(hip.HIPChain, pkg, L, U, beta, seed=3, delay_depth=W)
            chains.append(g)
            g.set_hopping(model.hopping_matrix())
            g.set_field(field.conf)
            g.set_seed(17)
            g.set_slice_launch(one)
            g.set_log(4 * 2 * len(model.l) * par.slices)
            g.build_stack()
        gs, gw = chains
        for it in range(4):
            if it == 2:
                for c in chains:       # a rebuild between sweeps resets the slice counters
                    c.build_stack()
            a1, a2 = gs.sweep(1), gw.sweep(1)
            assert a1 == a2, it
            assert gs.state() == gw.state(), it
            assert np.array_equal(gs.get_field(), gw.get_field()), it
            assert np.array_equal(gs.greens(), gw.greens()), it
        d1, p1 = gs.get_log()
        d2, p2 = gw.get_log()
        assert np.array_equal(d1, d2) and np.array_equal(p1, p2)
        assert gs.counters() == gw.counters()
        assert gs.counters()["device_errors"] == 0
        assert np.array_equal(gs.stabilization_log(), gw.stabilization_log())
    finally:
        for c in chains:
            c.close()


def test_slice_launch_toggles_mid_run(pkg, hip):
    """Switching the launch form between sweeps (and so between graph captures) keeps the
    chain on the same trajectory as a chain that never switches."""
    chains = []
    try:
        for _ in range(2):
            g, model, par, field = make_chain(hip.HIPChain, pkg, 8, 4.0, 2.0, seed=9)
            chains.append(g)
            g.set_hopping(model.hopping_matrix())
            g.set_field(field.conf)
            g.set_seed(5)
            g.build_stack()
        ga, gb = chains
        for it, mode in enumerate([True, False, True, True, False]):
            ga.set_slice_launch(mode)
            assert ga.sweep(1) == gb.sweep(1), it
            assert np.array_equal(ga.greens(), gb.greens()), it
        assert ga.counters()["device_errors"] == 0
    finally:
        for c in chains:
            c.close()
