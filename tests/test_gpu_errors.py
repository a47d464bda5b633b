"""Failure reporting of the HIP library: a kernel's bounded wait that gives up, and the
mixed-precision bound on max|G_wrapped - G_scratch| at every stabilisation.

Both conditions are forced or measured from the host; every kernel still ends normally."""
import numpy as np
import pytest

from conftest import make_chain

pytestmark = pytest.mark.gpu


def _chain(pkg, hip, L, U, beta, mu=0.0, **kw):
    g, model, par, field = make_chain(hip.HIPChain, pkg, L, U, beta, mu=mu, seed=17, **kw)
    g.set_hopping(model.hopping_matrix())
    g.set_field(field.conf)
    g.set_seed(5)
    return g, model, par, field


def test_window_wait_give_up_is_reported(pkg, hip):
    """A zero bound on the factor waves' wait makes them leave before the decisions are
    published: the launch ends, the chain counts a device error and dqmc_sweep returns
    DQMC_E_DEVICE.  With the default bound restored and the stack rebuilt, the chain sweeps
    without error again."""
    g, model, par, field = _chain(pkg, hip, 4, 4.0, 2.0)
    try:
        g.build_stack()
        g.set_wait_limit(0)
        with pytest.raises(hip.DQMCError, match=r"\(-5\).*gave up"):
            g.sweep(1)
        st = g.stats()
        assert st["device_errors"] > 0
        assert st["sweeps"] == 1          # every launch of the sweep ran to its end
        g.set_wait_limit(1 << 22)
        g.build_stack()
        g.sweep(1)                        # no new device error
        assert g.stats()["device_errors"] == st["device_errors"]
        with pytest.raises(hip.DQMCError):
            g.set_wait_limit(-1)
    finally:
        g.close()


def test_mixed_precision_bound_is_enforced_at_config_c(pkg, hip):
    """Attractive 24x24 (N=576), mu=-0.5, beta=12 (M=120): the mixed form (FP32 wraps and
    window updates between FP64 stabilisations) checks max|G_wrapped - G_scratch| at every
    stabilisation against its tolerance (1e-6 by default): dqmc_sweep returns
    DQMC_E_PRECISION exactly when some stabilisation exceeds it, and the count matches the
    stabilisation log.  The FP64 chain on the same field stays within the bound."""
    g, model, par, field = _chain(pkg, hip, 24, 4.0, 12.0, mu=-0.5, precision="mixed")
    f, *_ = _chain(pkg, hip, 24, 4.0, 12.0, mu=-0.5)
    try:
        g.build_stack()
        g.stabilization_log()
        try:
            g.sweep(1)
            raised = False
        except hip.DQMCError as e:
            assert "(-6)" in str(e)
            raised = True
        log = g.stabilization_log()
        st = g.stats()
        assert len(log) == 2 * 12
        over = int(np.sum(~(log <= 1e-6)))
        print(f"mixed config C: max|dG| per stabilisation {log.max():.3e}, above 1e-6: {over}/{len(log)}")
        assert raised == (over > 0)
        assert st["precision_violations"] == over
        f.build_stack()
        f.stabilization_log()
        f.sweep(1)
        flog = f.stabilization_log()
        print(f"FP64 config C: max|dG| per stabilisation {flog.max():.3e}")
        assert len(flog) == 2 * 12 and flog.max() < 1e-6
        assert f.stats()["precision_violations"] == 0
    finally:
        g.close()
        f.close()


def test_mixed_precision_with_a_loose_tolerance_runs(pkg, hip):
    """The tolerance is the caller's: a mixed chain with 1e-1 sweeps without error and logs
    every stabilisation."""
    g, *_ = _chain(pkg, hip, 6, 4.0, 2.0, precision="mixed", mixed_tolerance=1e-1)
    try:
        g.build_stack()
        g.sweep(2)
        log = g.stabilization_log()
        assert len(log) > 0 and log.max() <= 1e-1
        assert g.stats()["precision_violations"] == 0
    finally:
        g.close()
