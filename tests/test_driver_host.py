"""Host logic of the run! mirror: exit codes, the low-acceptance abort, checkpoints, the
time-limit save, resume and replay (CPU only: the chain is the oracle or a stub)."""
import os
import time

import numpy as np
import pytest


def test_exit_codes_match_reference(pkg):
    # helpers.jl:17-22
    E = pkg.ExitCode
    assert (E.GENERIC_FAILURE, E.SUCCESS, E.CANCELLED_TIME_LIMIT, E.CANCELLED_LOW_ACCEPTANCE) == (-1, 0, 1, 2)


class StubChain:
    """Accepts `accepted[k]` flips in sweep k (then none)."""

    def __init__(self, n, flv, M, sm, dtau, alpha, accepted=()):
        self.n, self.M = n, M
        self.accepted = list(accepted)
        self.k = 0
        self.conf = np.ones((n, M), dtype=np.int8, order="F")
        self.upos = 0

    def set_hopping(self, T): pass
    def set_field(self, c): self.conf[:, :] = c
    def get_field(self): return self.conf.copy(order="F")
    def set_seed(self, seed, position=0): self.upos = position
    def build_stack(self): pass
    def counters(self): return {"uniforms_used": self.upos, "negative_prob": 0}
    def stabilization_log(self): return []
    def greens(self, physical=False): return np.zeros((1, self.n, self.n))

    def sweep(self, k=1):
        a = self.accepted[self.k] if self.k < len(self.accepted) else 0
        self.k += 1
        return a


def test_low_acceptance_abort_uses_cumulative_rate(pkg, tmp_path):
    # DQMC.jl:289-302 with scheduler.jl:291-299: after round(1/min_update_rate) sweeps, abort
    # when accepted / total of the update is below the rate; one good early sweep must not
    # disable the check (a per-sweep maximum would).
    model = pkg.HubbardModel(2, 2, U=1.0)
    n, M = 4, 10
    full = 2 * n * M
    mc = pkg.DQMC(model, beta=1.0, thermalization=50, sweeps=0, seed=3, silent=True,
                  chain_factory=lambda *a: StubChain(*a, accepted=[full]))
    fail = str(tmp_path / "failed.npz")
    code = mc.run(verbose=False, min_update_rate=0.25, fail_filename=fail)
    # rates 1, 0, 0, 0, 0: the cumulative rate 1/5 < 0.25 first holds at sweep 5 > round(4)
    assert code == pkg.ExitCode.CANCELLED_LOW_ACCEPTANCE
    assert mc.last_sweep == 5
    assert os.path.isfile(fail)
    # a rate that stays above the threshold never aborts
    mc2 = pkg.DQMC(model, beta=1.0, thermalization=12, sweeps=0, seed=3, silent=True,
                   chain_factory=lambda *a: StubChain(*a, accepted=[full // 2] * 100))
    assert mc2.run(verbose=False, min_update_rate=0.25, fail_filename=fail) == pkg.ExitCode.SUCCESS
    assert mc2.last_sweep == 12


def _oracle_mc(pkg, oracle, **kw):
    model = pkg.HubbardModel(2, 2, U=4.0)
    factory = lambda n, f, M, sm, dt, a: oracle.OracleChain(n, f, M, sm, dt, a)
    return pkg.DQMC(model, beta=2.0, seed=11, silent=True, chain_factory=factory,
                    measure_rate=2, **kw)


def test_checkpoint_round_trip_and_resume_equals_uninterrupted_run(pkg, oracle, tmp_path):
    from montecarlo_jl_amd import fileio
    ref = _oracle_mc(pkg, oracle, thermalization=4, sweeps=8)
    assert ref.run(verbose=False) == pkg.ExitCode.SUCCESS

    # the same run, stopped by its time limit after the first sweep (safe_before is already
    # within the grace period), then resumed from the file
    f = str(tmp_path / "resumable.npz")
    mc = _oracle_mc(pkg, oracle, thermalization=4, sweeps=8)
    code = mc.run(verbose=False, safe_before=time.time(), grace_period=60.0, resumable_filename=f)
    assert code == pkg.ExitCode.CANCELLED_TIME_LIMIT
    assert mc.last_sweep == 1 and os.path.isfile(f)

    back = fileio.load(f, chain_factory=lambda n, fl, M, sm, dt, a: oracle.OracleChain(n, fl, M, sm, dt, a))
    assert back.last_sweep == 1
    assert np.array_equal(back.field.conf, mc.field.conf)
    assert back.rng_position == mc.rng_position > 0
    assert back.stream_seed == mc.stream_seed
    assert back.acceptance.total == mc.acceptance.total
    assert back.acceptance.accepted == mc.acceptance.accepted
    assert back.parameters == mc.parameters
    assert back.model.U == mc.model.U and len(back.model.l) == len(mc.model.l)

    code = back.run(verbose=False)
    assert code == pkg.ExitCode.SUCCESS
    assert back.last_sweep == ref.last_sweep == 12
    # the resumed chain rebuilds its stack from the field (G equal to rounding), and the
    # uniform stream continues from the stored position: same field, same recorded
    # configurations, same acceptance as the uninterrupted run
    assert np.array_equal(back.field.conf, ref.field.conf)
    assert len(back.recorder) == len(ref.recorder) == 4
    for a, b in zip(back.recorder, ref.recorder):
        assert np.array_equal(a, b)
    assert back.acceptance.accepted == pytest.approx(ref.acceptance.accepted, abs=0)
    m_ref, m_back = ref.measurements["equal_time"], back.measurements["equal_time"]
    assert m_back.count == m_ref.count
    np.testing.assert_allclose(m_back.mean("density"), m_ref.mean("density"), rtol=1e-9)


def test_save_rename_and_overwrite(pkg, oracle, tmp_path):
    from montecarlo_jl_amd import fileio
    mc = _oracle_mc(pkg, oracle, thermalization=1, sweeps=0)
    mc.run(verbose=False)
    f = str(tmp_path / "state.npz")
    assert fileio.save(f, mc) == f
    f2 = fileio.save(f, mc)                  # rename = true: a new unique name
    assert f2 != f and os.path.isfile(f2)
    with pytest.raises(FileExistsError):
        fileio.save(f, mc, rename=False)
    assert fileio.save(f, mc, overwrite=True) == f
    assert not [p for p in os.listdir(tmp_path) if p.startswith(".")]   # backup removed
    with np.load(f, allow_pickle=False) as z:
        assert int(z["VERSION"]) == 2 and str(z["MC/tag"]) == "DQMC"
        assert str(z["MC/field/name"]) == "DensityHirschField"
        # the field in BitArray chunks, as Julia stores compress(field)
        conf = mc.field.conf
        assert np.array_equal(z["MC/field/conf"], pkg.fields.bits_to_chunks(conf == 1))


def test_replay_measures_recorded_configurations(pkg, oracle):
    mc = _oracle_mc(pkg, oracle, thermalization=2, sweeps=6)
    mc.run(verbose=False)
    configs = list(mc.recorder)
    assert len(configs) == 3
    # one configuration: the sample is the physical G of that field computed from scratch at
    # slice 0 (calculate_greens(mc, 0), DQMC.jl:465)
    one = _oracle_mc(pkg, oracle, thermalization=2, sweeps=6)
    assert one.replay(configs[:1]) == pkg.ExitCode.SUCCESS
    probe = _oracle_mc(pkg, oracle, thermalization=2, sweeps=6)
    probe.init()
    probe.field.decompress(pkg.fields.chunks_to_bits(configs[0], probe.field.conf.shape))
    probe.stack.set_field(probe.field.conf)
    probe.stack.build_stack()
    G0 = probe.stack.calculate_greens(0)
    eTmh, eTph, _, _ = probe.stack.get_hopping_exp()
    ms = pkg.measurements.MeasurementSet(mc.model.hopping_matrix(), mc.model.U)
    ms.push(np.stack([eTph @ g @ eTmh for g in G0]))
    got = one.measurements["equal_time"]
    assert got.count == 1
    for key in ("density", "kinetic_energy", "interaction_energy"):
        np.testing.assert_allclose(got.mean(key), ms.mean(key), rtol=1e-12, atol=1e-14)
    # all of them
    allc = _oracle_mc(pkg, oracle, thermalization=2, sweeps=6)
    assert allc.replay(configs) == pkg.ExitCode.SUCCESS
    assert allc.measurements["equal_time"].count == 3
