import importlib.util
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "montecarlo.jl_amd")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def load_pkg():
    """Import montecarlo.jl_amd (its directory name is not a Python identifier)."""
    name = "montecarlo_jl_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def pkg():
    return load_pkg()


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.build()
    return O


@pytest.fixture(scope="session")
def hip(pkg):
    from montecarlo_jl_amd import _lib
    _lib.load()
    return _lib


def make_chain(cls, pkg, L, U_ref, beta, *, mu=0.0, lattice=None, safe_mult=10, seed=1234,
               delta_tau=0.1, **kw):
    """A chain of the reference's HubbardModel(L, 2; U, mu) with a seeded +-1 field."""
    model = pkg.HubbardModel(L, 2, U=U_ref, mu=mu) if lattice is None else \
        pkg.HubbardModel(l=lattice, U=U_ref, mu=mu)
    par = pkg.DQMCParameters.create(beta=beta, delta_tau=delta_tau, safe_mult=safe_mult)
    field = model.choose_field()(par, model)
    field.rand(seed)
    ch = cls(len(model.l), field.flavors, par.slices, safe_mult, par.delta_tau, field.alpha, **kw)
    return ch, model, par, field


def rel_maxdiff(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


TIE_BAND = 1e-12


def tie_positions(p: np.ndarray, stream: np.ndarray) -> np.ndarray:
    """Proposals whose Metropolis test `p > 1 || u < p` (local_updates.jl:53) lies within the
    tie band: |u - p| < 1e-12 when a uniform is drawn (p <= 1), 1 < p < 1 + 1e-12 otherwise.
    `stream` holds the uniforms in the order they are consumed (one per proposal with p <= 1)."""
    p = np.asarray(p, dtype=np.float64)
    need = ~(p > 1.0)
    idx = np.clip(np.cumsum(need) - 1, 0, max(len(stream) - 1, 0))
    u = np.asarray(stream, dtype=np.float64)[idx] if len(stream) else np.ones_like(p)
    tie = np.where(need, np.abs(u - p) < TIE_BAND, p < 1.0 + TIE_BAND)
    return np.flatnonzero(tie)


def assert_same_decisions(dec_gpu, dec_oracle, p_oracle, stream) -> bool:
    """The accept sequences agree bit for bit up to the first tie (where the two sides may
    legitimately part); returns True when there is no tie at all."""
    ties = tie_positions(p_oracle, stream)
    upto = int(ties[0]) if len(ties) else len(dec_oracle)
    assert len(dec_gpu) == len(dec_oracle)
    mism = np.flatnonzero(dec_gpu[:upto] != dec_oracle[:upto])
    assert len(mism) == 0, f"decisions differ at {mism[:5]} before the first tie ({upto})"
    return len(ties) == 0


def counter_stream(seed: int, count: int) -> np.ndarray:
    """The library's counter stream u_k = SplitMix64(seed, k) >> 11 * 2^-53 from position 0."""
    from montecarlo_jl_amd.fields import uniforms
    return uniforms(seed, 0, count)
