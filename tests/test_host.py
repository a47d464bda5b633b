"""Host logic (no GPU): lattices, hopping matrix, parameters, fields, RNG streams, C-ABI exports."""
import json
import math
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT

GOLDEN = os.path.join(ROOT, "tests", "golden")


def test_directed_bonds_match_reference_known_answers(pkg):
    """test/lattices.jl:87-88 (Square(3)) and :176-177 (Honeycomb(2))."""
    ref = json.load(open(os.path.join(GOLDEN, "lattice_bonds.json")))
    for lat, key in ((pkg.SquareLattice(3), "square3_directed"), (pkg.Honeycomb(2), "honeycomb2_directed")):
        b = lat.bonds(directed=True)
        assert [f + 1 for f, _ in b] == ref[key]["from"]
        assert [t + 1 for _, t in b] == ref[key]["to"]


@pytest.mark.parametrize("lat,coord", [("square", 4), ("honeycomb", 3), ("chain", 2), ("cubic", 6)])
def test_hopping_matrix_invariants(pkg, lat, coord):
    L = {"square": pkg.SquareLattice(5), "honeycomb": pkg.Honeycomb(4), "chain": pkg.Chain(7),
         "cubic": pkg.CubicLattice(3)}[lat]
    m = pkg.HubbardModel(l=L, U=2.0, mu=0.3, t=1.3)
    T = m.hopping_matrix()
    assert np.array_equal(T, T.T)
    assert np.allclose(np.diag(T), -0.3)
    assert np.allclose(T.sum(axis=1), -coord * 1.3 - 0.3)


def test_hopping_matrix_L2_bond_doubling(pkg):
    """SquareLattice(2): +x and -x reach the same neighbour, so T[i,j] = -2t (HubbardModel.jl:112-124)."""
    T = pkg.HubbardModel(2, 2, U=1.0, t=1.0).hopping_matrix()
    assert T[0, 1] == -2.0 and T[0, 2] == -2.0 and T[0, 3] == 0.0


def test_parameters(pkg):
    """parameters.jl:84-139 and the checks of test/flavortests_DQMC.jl:4-18."""
    p = pkg.DQMCParameters.create(beta=5.0)
    assert p.delta_tau == 0.1 and p.slices == 50
    p = pkg.DQMCParameters.create(beta=5.0, slices=40)
    assert math.isclose(p.delta_tau, 0.125)
    p = pkg.DQMCParameters.create(delta_tau=0.1, slices=50)
    assert math.isclose(p.beta, 5.0)
    with pytest.raises(ValueError):
        pkg.DQMCParameters.create(beta=5.0, delta_tau=0.1, slices=49)
    with pytest.raises(ValueError):
        pkg.DQMCParameters.create(beta=5.0, foo=1)


def test_generate_chunks(pkg, oracle):
    """stack.jl:154-158; invariants of test/flavortests_DQMC.jl:245-262; oracle agreement."""
    rng = np.random.default_rng(0)
    for _ in range(200):
        M, s = int(rng.integers(1, 101)), int(rng.integers(1, 13))
        ch = pkg.generate_chunks(M, s)
        assert ch[0][0] == 1 and ch[-1][1] == M
        assert all(a <= b and b - a + 1 <= s for a, b in ch)
        assert all(ch[i][1] + 1 == ch[i + 1][0] for i in range(len(ch) - 1))
        assert ch == oracle.generate_chunks(M, s)
    assert pkg.generate_chunks(40, 10) == [(1, 10), (11, 20), (21, 30), (31, 40)]
    assert pkg.generate_chunks(15, 10) == [(1, 8), (9, 15)]   # Julia rounds 7.5 to even


def test_hirsch_fields(pkg):
    """fields.jl:363-376, 412-425; choose_field HubbardModel.jl:83; Repulsive negates U."""
    par = pkg.DQMCParameters.create(beta=1.0)
    att = pkg.AttractiveHubbardModel(4, 2, U=4.0)
    rep = pkg.RepulsiveHubbardModel(4, 2, U=4.0)
    assert rep.U == -4.0
    fa = att.choose_field()(par, att)
    fr = rep.choose_field()(par, rep)
    assert isinstance(fa, pkg.DensityHirschField) and fa.flavors == 1
    assert isinstance(fr, pkg.MagneticHirschField) and fr.flavors == 2
    a = math.acosh(math.exp(0.5 * 0.1 * 4.0))
    assert math.isclose(fa.alpha, a, rel_tol=1e-15) and math.isclose(fr.alpha, a, rel_tol=1e-15)
    assert math.isclose(a, 0.6537326, rel_tol=1e-6)
    with pytest.raises(ValueError):
        pkg.MagneticHirschField(par, att)   # mismatched sign gives a complex coupling
    fa.rand(5)
    assert set(np.unique(fa.conf)) <= {-1, 1} and fa.conf.dtype == np.int8
    assert fa.conf.shape == (16, 10) and fa.conf.flags.f_contiguous
    c = fa.compress()
    fa2 = pkg.DensityHirschField(par, att)
    fa2.decompress(c)
    assert np.array_equal(fa.conf, fa2.conf)


def test_field_chunks_match_julia_bitarray_layout(pkg):
    """compress(f) = BitArray(conf .== 1) (fields.jl:331) in Julia's storage: column-major
    entry k at bit k % 64 of uint64 chunk k // 64, unused high bits zero."""
    from montecarlo_jl_amd.fields import bits_to_chunks, chunks_to_bits
    conf = np.array([[1, -1], [-1, -1], [1, 1]], dtype=np.int8, order="F")   # entries 0, 2, 5 are +1
    assert bits_to_chunks(conf == 1).tolist() == [1 + 4 + 32]
    b = np.zeros(65, dtype=bool)
    b[63] = b[64] = True
    assert bits_to_chunks(b).tolist() == [1 << 63, 1]
    # against numpy's little-endian bit packing (a library routine), ragged lengths
    rng = np.random.default_rng(3)
    for shape in [(16, 10), (36, 15), (7, 9), (256, 100), (1, 1)]:
        bits = rng.random(shape) < 0.5
        flat = bits.ravel(order="F")
        nch = (flat.size + 63) // 64
        by = np.zeros(nch * 8, dtype=np.uint8)
        pk = np.packbits(flat, bitorder="little")
        by[:pk.size] = pk
        assert np.array_equal(bits_to_chunks(bits), by.view("<u8"))
        assert np.array_equal(chunks_to_bits(bits_to_chunks(bits), shape), bits)
    par = pkg.DQMCParameters.create(beta=1.5, delta_tau=0.1)
    model = pkg.AttractiveHubbardModel(6, 2, U=4.0)
    f = pkg.DensityHirschField(par, model)
    f.rand(9)
    c = f.compress_chunks()
    g = pkg.DensityHirschField(par, model)
    g.decompress_chunks(c)
    assert c.size == (36 * 15 + 63) // 64 and np.array_equal(f.conf, g.conf)


def test_uniform_stream_matches_oracle(oracle):
    from montecarlo_jl_amd.fields import uniforms
    u = uniforms(1234, 10, 50)
    ref = np.array([oracle.lib().orc_uniform_at(1234, k) for k in range(10, 60)])
    assert np.array_equal(u, ref)
    assert (u >= 0).all() and (u < 1).all()


def test_measurement_kernels_free_fermions(pkg):
    """At U=0 the equal-time kernels reproduce the exact free-fermion expectation values."""
    from montecarlo_jl_amd import measurements as ms
    m = pkg.HubbardModel(3, 2, U=0.0, mu=0.2)
    T = m.hopping_matrix()
    w, V = np.linalg.eigh(T)
    beta = 2.0
    G = V @ np.diag(1 / (1 + np.exp(-beta * w))) @ V.T
    nk = 1 / (1 + np.exp(beta * w))
    assert math.isclose(ms.kinetic_energy(G[None], T), 2 * float(np.sum(w * nk)), rel_tol=1e-12)
    assert math.isclose(ms.density(G[None]), 2 * float(np.sum(nk)) / len(w), rel_tol=1e-12)


def test_c_abi_exports_every_declared_symbol(pkg):
    """The library loads without a GPU and exports every function include/dqmc_hip.h declares."""
    from montecarlo_jl_amd import _lib
    header = open(os.path.join(ROOT, "include", "dqmc_hip.h")).read()
    declared = set(re.findall(r"^\s*(?:int|const char\*)\s+(dqmc_\w+)\s*\(", header, re.M))
    assert declared and declared == set(_lib.EXPORTED)
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\bT (dqmc_\w+)", out))
    assert declared <= exported, declared - exported
    L = _lib.load()
    for name in declared:
        assert hasattr(L, name)


def test_footprint_is_monotone(pkg):
    from montecarlo_jl_amd import _lib
    a = _lib.footprint_bytes(256, 2, 100, 10)
    b = _lib.footprint_bytes(576, 1, 120, 10)
    c = _lib.footprint_bytes(256, 2, 100, 10, "mixed")
    assert 20e6 < a < 80e6 and b > a and c > a


def test_c_harness_is_built_and_loads_the_library():
    """The compiled C harness (tests/c_harness) links against the in-tree library and starts
    without a GPU (usage message, exit status 2, before any library call)."""
    exe = os.path.join(ROOT, "tests", "c_harness", "dqmc_harness")
    if not os.path.exists(exe):
        subprocess.run(["make", "-C", os.path.dirname(exe)], check=True, capture_output=True)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr
