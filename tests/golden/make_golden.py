"""Writes the committed fixtures of tests/golden/.

  lattice_bonds.json  directed bond lists of Square(3) and Honeycomb(2), transcribed from the
                      reference's known-answer tests test/lattices.jl:87-88 and :176-177
                      (1-based, as there)
  config_a.npz        CPU-oracle vectors for the attractive 4x4 model (U=4 reference
                      convention, beta=4, dtau=0.1, M=40): seeded field, effective G at
                      slices 0/13/40, a uniform stream with the accept bitmap, acceptance
                      probabilities and final field of one sweep
  ed_2x2.json         exact-diagonalisation values for the reference's finite-U ED test
                      models (test/ED/ED_tests.jl:402-406): 2x2, beta=1

Run from the repository root: python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

SQUARE3_FROM = [1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 6, 6, 6, 6,
                7, 7, 7, 7, 8, 8, 8, 8, 9, 9, 9, 9]
SQUARE3_TO = [2, 4, 3, 7, 3, 5, 1, 8, 1, 6, 2, 9, 5, 7, 6, 1, 6, 8, 4, 2, 4, 9, 5, 3,
              8, 1, 9, 4, 9, 2, 7, 5, 7, 3, 8, 6]
HONEYCOMB2_FROM = [1, 1, 1, 5, 5, 5, 2, 2, 2, 6, 6, 6, 3, 3, 3, 7, 7, 7, 4, 4, 4, 8, 8, 8]
HONEYCOMB2_TO = [5, 6, 7, 1, 2, 3, 6, 5, 8, 2, 1, 4, 7, 8, 5, 3, 4, 1, 8, 7, 6, 4, 3, 2]


def write_config_a_raw(T):
    """config_a_raw.bin: the config_a.npz vectors for the C harness (tests/c_harness), raw
    little-endian: int32 n, slices, count = 2 n slices; float64 alpha; float64 T[n*n],
    G_start[n*n], G_end[n*n], G_slice0[n*n], G_slice13[n*n], G_slice40[n*n] (column-major);
    int8 conf[n*slices], final_conf[n*slices]; float64 uniforms[count], p[count]; uint8
    accepted[count]."""
    z = np.load(os.path.join(HERE, "config_a.npz"))
    n, M = z["conf"].shape
    cnt = 2 * n * M
    with open(os.path.join(HERE, "config_a_raw.bin"), "wb") as f:
        f.write(np.array([n, M, cnt], dtype="<i4").tobytes())
        f.write(np.array([float(z["alpha"])], dtype="<f8").tobytes())
        for key in ("G_start", "G_end", "G_slice0", "G_slice13", "G_slice40"):
            if key == "G_start":
                f.write(np.asarray(T, dtype="<f8").ravel(order="F").tobytes())
            f.write(np.asarray(z[key], dtype="<f8").ravel(order="F").tobytes())
        for key in ("conf", "final_conf"):
            f.write(np.asarray(z[key], dtype=np.int8).ravel(order="F").tobytes())
        f.write(np.asarray(z["uniforms"], dtype="<f8").tobytes())
        f.write(np.asarray(z["p"], dtype="<f8").tobytes())
        f.write(np.asarray(z["accepted"], dtype=np.uint8).tobytes())


def main():
    from conftest import load_pkg
    from oracle import oracle as O
    from oracle.ed import HubbardED
    pkg = load_pkg()
    O.build()

    with open(os.path.join(HERE, "lattice_bonds.json"), "w") as f:
        json.dump({"square3_directed": {"from": SQUARE3_FROM, "to": SQUARE3_TO},
                   "honeycomb2_directed": {"from": HONEYCOMB2_FROM, "to": HONEYCOMB2_TO},
                   "source": "test/lattices.jl:87-88, 176-177 (1-based)"}, f, indent=1)

    model = pkg.HubbardModel(4, 2, U=4.0)
    par = pkg.DQMCParameters.create(beta=4.0, delta_tau=0.1, safe_mult=10)
    field = model.choose_field()(par, model)
    field.rand(2024)
    n, M = len(model.l), par.slices
    ch = O.OracleChain(n, 1, M, 10, par.delta_tau, field.alpha)
    ch.set_hopping(model.hopping_matrix())
    ch.set_field(field.conf)
    g = {f"G_slice{s}": ch.calculate_greens(s)[0] for s in (0, 13, M)}
    u = np.random.default_rng(31).random(2 * n * M)
    ch.set_uniforms(u)
    ch.set_log(2 * n * M)
    ch.build_stack()
    G_start = ch.greens()[0]
    ch.sweep(1)
    dec, p = ch.get_log()
    np.savez_compressed(os.path.join(HERE, "config_a.npz"), conf=field.conf, uniforms=u,
                        accepted=dec, p=p, final_conf=ch.get_field(), G_start=G_start,
                        G_end=ch.greens()[0], alpha=field.alpha, **g)
    write_config_a_raw(model.hopping_matrix())

    ed = {}
    for name, U, mu in (("repulsive", -1.0, 0.0), ("attractive_mu1", 1.0, 1.0)):
        m = pkg.HubbardModel(2, 2, U=U, mu=mu, t=1.0)
        e = HubbardED(m, beta=1.0)
        ed[name] = {"U": U, "mu": mu, "energy": e.energy(), "greens": e.greens().tolist(),
                    "occupation": e.occupation().tolist(),
                    "double_occupancy": e.double_occupancy().tolist()}
    with open(os.path.join(HERE, "ed_2x2.json"), "w") as f:
        json.dump(ed, f, indent=1)


if __name__ == "__main__":
    main()
