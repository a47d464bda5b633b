#!/usr/bin/env python3
"""DQMC sweep benchmark on one MI355X.

Workload (one "step" = one sweep = one LocalSweep: 2M slice visits x N Metropolis proposals
including every wrap and stabilisation, no measurements):
  Repulsive Hubbard model, 16x16 periodic square lattice (N=256), t=1, U=4, mu=0, beta=10,
  dtau=0.1 (M=100), magnetic Hirsch field (2 flavours, block diagonal), safe_mult=10,
  one chain, seeded +-1 field, counter-stream uniforms.  FP64 throughout (headline) and the
  mixed-precision form (FP32 wraps and rank-1 updates, FP64 UDT/recompute/detratio).

Prints ONE JSON line.  Inputs are resident on the device before the timed region.
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))

# The batched and serving legs run one stream per chain: give the process 16 hardware queues
# (HIP's default is 4) so that their streams do not share queues.  Set before the first HIP
# call; a value from the environment wins.  The one-chain headline uses a single stream.
os.environ.setdefault("GPU_MAX_HW_QUEUES", "16")
METRIC = "DQMC sweeps/sec (N=256, M=100) + p50/p99 chain latency, mixed vs FP64"
PEAK = {"fp64_mfma_tflops": 78.6, "fp32_mfma_tflops": 157.3, "hbm_gbs": 8000.0}


def load_pkg():
    name = "montecarlo_jl_amd"
    if name in sys.modules:
        return sys.modules[name]
    d = os.path.join(ROOT, "montecarlo.jl_amd")
    spec = importlib.util.spec_from_file_location(name, os.path.join(d, "__init__.py"),
                                                  submodule_search_locations=[d])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def config_b(pkg):
    model = pkg.RepulsiveHubbardModel(16, 2, U=4.0)   # reference convention: U_ref = -4
    par = pkg.DQMCParameters.create(beta=10.0, delta_tau=0.1, safe_mult=10)
    field = model.choose_field()(par, model)
    field.rand(1)
    return model, par, field


def config_c(pkg):
    """Attractive 24x24 (N=576), U=4 in the reference's convention (attractive), mu=-0.5,
    beta=12, dtau=0.1 (M=120), density field (1 flavour)."""
    model = pkg.AttractiveHubbardModel(24, 2, U=4.0, mu=-0.5)
    par = pkg.DQMCParameters.create(beta=12.0, delta_tau=0.1, safe_mult=10)
    field = model.choose_field()(par, model)
    field.rand(3)
    return model, par, field


def make_chain(pkg, model, par, field, precision):
    from montecarlo_jl_amd import _lib
    ch = _lib.HIPChain(len(model.l), field.flavors, par.slices, par.safe_mult, par.delta_tau,
                       field.alpha, check_propagation_error=True, precision=precision)
    ch.set_hopping(model.hopping_matrix())
    ch.set_field(field.conf)
    ch.set_seed(12345)
    ch.build_stack()
    return ch


def timed_sweeps(ch, steps, warmup):
    for _ in range(warmup):
        ch.sweep(1)
    ch.synchronize()
    lat = []
    acc0 = ch.stats()["accepted"]
    t0 = time.perf_counter()
    for _ in range(steps):
        a = time.perf_counter()
        ch.sweep(1)            # returns after the sweep's device work is complete
        lat.append(time.perf_counter() - a)
    ch.synchronize()
    total = time.perf_counter() - t0
    acc = ch.stats()["accepted"] - acc0
    lat = np.array(lat) * 1e3
    return total, lat, acc


def udt_panels(n):
    """Panel starts and widths of the blocked UDT (qr.hip: udt_panel_width)."""
    out, p0 = [], 0
    while p0 < n:
        rows, rpt = n - p0, 1
        while rpt * 256 < rows:
            rpt *= 2
        w = max(1, min(64 // rpt, 16, rows))
        out.append((p0, w))
        p0 += w
    return out


def kernel_work(kind, n, flv, win, per_launch_accepted=0.0, windows_per_launch=1):
    """Algorithmic FLOP per launch of each kernel kind at block size n with flv blocks.

    gemm: one n^3 product per block.  udt_update: block-reflector updates, 4 rows w cols
    flops per strip set, averaged over the trailing and the Q launches of one UDT.
    udt_panel: Householder steps on the rows x w panel.  rdivp: triangular solve with n
    right-hand rows.  local_apply: G - C (Y R) with m accepted flips per window.  local:
    rank-1 updates of the window block plus the two m x m triangular solves.
    """
    if kind in ("gemm_f64", "gemm_f32"):
        return 2.0 * flv * n ** 3
    if kind == "rdivp":
        return flv * float(n) ** 3
    if kind == "udt_update":
        pans = udt_panels(n)
        tot, cnt = 0.0, 0
        for p0, w in pans:
            rows = n - p0
            if p0 + w < n:
                tot += 4.0 * rows * w * (n - p0 - w)
                cnt += 1
            tot += 4.0 * rows * w * rows
            cnt += 1
        return flv * tot / cnt
    if kind == "udt_panel":
        pans = udt_panels(n)
        return flv * sum(2.0 * (n - p0) * w * w for p0, w in pans) / len(pans)
    nw = -(-n // win)                 # windows per slice; the last one's update runs alone
    if kind == "local_apply":
        m = per_launch_accepted
        return 2.0 * flv * (n * n * m + m * m * n)
    if kind == "local":
        # window decisions and window updates of G: a launch per window carries its own
        # window and the previous window's update (all but the first of a slice); a launch
        # per slice carries the slice's windows and every update
        m = per_launch_accepted / windows_per_launch
        own = flv * (2.0 * win * win * m + (2.0 / 3.0) * m ** 3)
        upd = 2.0 * flv * (n * n * m + m * m * n)
        if windows_per_launch == 1:
            return own + (nw - 1) / nw * upd
        return windows_per_launch * (own + upd)
    return 0.0


def kernel_bytes(kind, n, flv, win, windows_per_launch=1):
    """Algorithmic HBM bytes per launch of the window kernel: per window the window block,
    the strips G[w, wp], G[wp, w] and the previous factor read, the factor written, field
    column and log; per window update G read once and written once (in all but the first
    window launch of a slice, or every window of a slice launch)."""
    if kind == "local":
        nw = -(-n // win)
        own = 4.0 * flv * win * win * 8 + flv * win * win * 8 + 3 * win + 4
        upd = 2.0 * flv * n * n * 8
        if windows_per_launch == 1:
            return own + (nw - 1) / nw * upd
        return windows_per_launch * (own + upd)
    return 0.0


def rocprof_avg_us(kernel_tag):
    """Average duration (us) of a kernel in the committed rocprofv3 --kernel-trace --stats
    summary of the config-B sweep (profiles/dqmc_sweep_config_b_kernel_stats.md), or None."""
    path = os.path.join(ROOT, "profiles", "dqmc_sweep_config_b_kernel_stats.md")
    try:
        for line in open(path):
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            if len(cells) >= 4 and kernel_tag in cells[0]:
                return float(cells[3])
    except (OSError, ValueError):
        pass
    return None


def measured_traffic(kernel_tag):
    """HBM bytes per launch of a kernel from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    summary (profiles/dqmc_sweep_config_b_hbm_traffic.md), or None."""
    path = os.path.join(ROOT, "profiles", "dqmc_sweep_config_b_hbm_traffic.md")
    try:
        for line in open(path):
            if line.startswith(kernel_tag + " traffic_bytes_per_launch:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


# servers whose requests outlived the serving leg's deadline: kept open (closing one waits for
# its running requests), and the process leaves with os._exit once its line is printed
_LEFT_RUNNING = []


def serve_leg(pkg, nreq, rate=20.0, seed=2024, sweep_cap=2, max_wall=150.0):
    """Config E's request mix through the serving loop: Poisson arrivals at `rate` per second
    of chains with L uniform in {8..32}, square (N = L^2) or honeycomb (N = 2 L^2) lattice,
    beta uniform in [5, 20] (M = round(10 beta)), U uniform in [-8, 8], dtau = 0.1, FP64.
    The drawn sweep count (uniform in {200..2000}) is capped at `sweep_cap` per request so
    that the leg takes about a minute (a full request runs for up to hours); shapes, beta and
    U are drawn as in the full mix.  Admission by HBM footprint (64 GiB budget), at most 16
    chains at once; latency is submit -> result (stack build included).  Requests still
    running `max_wall` seconds after the first submission are reported, not waited for."""
    from montecarlo_jl_amd import _lib as L
    rng = np.random.default_rng(seed)
    srv = L.Server(memory_budget=64 << 30, max_active=16)
    t0 = time.perf_counter()
    ids, sweeps, t_next, drawn, shapes = [], 0, 0.0, 0, []
    pending = []
    try:
        for k in range(nreq):
            t_next += rng.exponential(1.0 / rate)
            while time.perf_counter() - t0 < t_next:
                time.sleep(0.001)
            Lx = int(rng.integers(8, 33))
            lat = pkg.SquareLattice(Lx) if rng.random() < 0.5 else pkg.Honeycomb(Lx)
            U = float(rng.uniform(-8.0, 8.0)) or 1.0
            beta = float(rng.uniform(5.0, 20.0))
            ns_full = int(rng.integers(200, 2001))
            ns = min(ns_full, sweep_cap)
            model = pkg.HubbardModel(l=lat, U=U, mu=0.0)
            par = pkg.DQMCParameters.create(beta=beta, delta_tau=0.1, safe_mult=10, warn_round=False)
            field = model.choose_field()(par, model)
            field.rand(1000 + k)
            ids.append(srv.submit(model.hopping_matrix(), field.conf, field.flavors, par.safe_mult,
                                  par.delta_tau, field.alpha, 5000 + k, ns))
            sweeps += ns
            drawn += ns_full
            shapes.append(len(lat))
        lats, queued, failed = [], [], 0
        pending = list(range(len(ids)))
        while pending and time.perf_counter() - t0 < max_wall:
            for i in list(pending):
                st = srv.poll(ids[i])
                if st not in ("done", "failed"):
                    continue
                pending.remove(i)
                if st == "done":
                    r = srv.result(ids[i])
                    lats.append(r["latency_s"])
                    queued.append(r["queued_s"])
                else:
                    failed += 1
                progress(f"serving: request {i} ({shapes[i]} sites) {st} at "
                         f"{time.perf_counter() - t0:.1f} s")
            time.sleep(0.02)
        wall = time.perf_counter() - t0
        peak = srv.status()["peak_active"]
    finally:
        if pending:
            _LEFT_RUNNING.append(srv)
        else:
            srv.close()
    if pending or failed or not lats:
        return {"requests": nreq, "status": "incomplete", "completed": len(lats), "failed": failed,
                "still_running": len(pending), "wall_s": wall, "max_wall_s": max_wall,
                "still_running_sites": [shapes[i] for i in pending]}
    return {"requests": nreq, "arrival_rate_per_s": rate, "sweeps": sweeps, "wall_s": wall,
            "aggregate_sweeps_per_s": sweeps / wall, "peak_active": peak,
            "latency_ms": {"p50": float(np.percentile(lats, 50)) * 1e3,
                           "p99": float(np.percentile(lats, 99)) * 1e3},
            "queued_ms_p99": float(np.percentile(queued, 99)) * 1e3,
            "sites": {"min": int(min(shapes)), "max": int(max(shapes)), "mean": float(np.mean(shapes))},
            "sweep_cap": sweep_cap, "sweeps_drawn": drawn,
            "workload": "config E mix: square or honeycomb L in {8..32}, beta U[5,20], U U[-8,8], "
                        f"FP64; sweeps per request capped at {sweep_cap} (drawn from {{200..2000}})"}


def config_d_specs(pkg, count, seed=2024):
    """Config D: L uniform in {6, 8, 10, 12, 16} (square), U uniform in [-8, 8] (reference
    convention: negative = repulsive, two flavours), beta in {4, 6, 8, 10}, dtau = 0.1, mu = 0,
    a seeded +-1 field per chain."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        L = int(rng.choice([6, 8, 10, 12, 16]))
        U = float(rng.uniform(-8.0, 8.0)) or 1.0
        beta = float(rng.choice([4.0, 6.0, 8.0, 10.0]))
        model = pkg.HubbardModel(L, 2, U=U)
        par = pkg.DQMCParameters.create(beta=beta, delta_tau=0.1, safe_mult=10)
        field = model.choose_field()(par, model)
        field.rand(100 + k)
        out.append((model, par, field, 500 + k))
    return out


def batched_leg(pkg, count, sweeps):
    """Config D: `count` independent FP64 chains swept together by dqmc_batch_sweep (lockstep
    slice visits on one stream, every launch covering all chains).  Each chain makes `sweeps`
    sweeps (2 M slice visits each); chains with fewer slices finish earlier.  Per-chain sweep
    latency: the chain's 2 M visits at the measured time per lockstep step."""
    from montecarlo_jl_amd import _lib as L
    specs = config_d_specs(pkg, count)
    chains = []
    for model, par, field, seed in specs:
        c = L.HIPChain(len(model.l), field.flavors, par.slices, par.safe_mult, par.delta_tau, field.alpha)
        c.set_hopping(model.hopping_matrix())
        c.set_field(field.conf)
        c.set_seed(seed)
        c.build_stack()
        chains.append(c)
    try:
        L.batch_sweep(chains, 1)                       # warm-up
        t0 = time.perf_counter()
        L.batch_sweep(chains, sweeps)
        wall = time.perf_counter() - t0
    finally:
        for c in chains:
            c.close()
    Ms = np.array([p.slices for _, p, _, _ in specs])
    ns = np.array([len(m.l) for m, _, _, _ in specs])
    flv = np.array([f.flavors for _, _, f, _ in specs])
    step_s = wall / (2 * Ms.max() * sweeps)
    lat = 2 * Ms * step_s
    K = -(-Ms // 10)
    flop = float(np.sum(flv * ns.astype(float) ** 3 * (12 * Ms + 46 * K))) * sweeps
    return {"workload": "config D: 64 chains, square L in {6,8,10,12,16}, U uniform [-8,8], "
                        "beta in {4,6,8,10}, dtau=0.1, mu=0, FP64, lockstep slice visits on one stream",
            "chains": count, "sweeps_per_chain": sweeps, "wall_s": wall,
            "aggregate_sweeps_per_s": count * sweeps / wall,
            "latency_ms": {"p50": float(np.percentile(lat, 50)) * 1e3, "p99": float(np.percentile(lat, 99)) * 1e3,
                           "basis": "2 M slice visits x measured time per lockstep step"},
            "step_us": step_s * 1e6,
            "dense_tflops": flop / wall / 1e12,
            "sites": {int(k): int(v) for k, v in zip(*np.unique(ns, return_counts=True))}}


def mixed_leg(pkg, sweeps, tol=1e-6):
    """Config C (N=576, M=120): the FP64 chain and the mixed-precision chain (FP32 wraps and
    window updates between FP64 stabilisations) from the same field and uniforms.  The mixed
    leg reports a rate only if every stabilisation's max|G_wrapped - G_scratch| stays within
    `tol`; otherwise it is marked failed with the measured drift."""
    from montecarlo_jl_amd import _lib as L
    model, par, field = config_c(pkg)
    out = {"workload": "attractive Hubbard 24x24 square, N=576, U=4 (attractive), mu=-0.5, "
                       "beta=12, dtau=0.1, M=120, density field, safe_mult=10, 1 chain",
           "tolerance": tol, "sweeps": sweeps}
    res = {}
    for prec in ("fp64", "mixed"):
        ch = make_chain(pkg, model, par, field, prec)
        ch.stabilization_log()
        lat, logs = [], []
        t0 = time.perf_counter()
        for _ in range(sweeps):
            a = time.perf_counter()
            try:
                ch.sweep(1)
            except L.DQMCError as e:
                if "(-6)" not in str(e):
                    raise
            lat.append(time.perf_counter() - a)
            logs.append(ch.stabilization_log())
        total = time.perf_counter() - t0
        dg = np.concatenate(logs)
        res[prec] = {"sweeps_per_s": sweeps / total, "p50_ms": float(np.percentile(lat, 50)) * 1e3,
                     "p99_ms": float(np.percentile(lat, 99)) * 1e3,
                     "max_dG_per_stabilisation": float(dg.max()), "stabilisations": int(len(dg)),
                     "stabilisations_above_tolerance": int(np.sum(~(dg <= tol)))}
        ch.close()
    out["fp64"] = res["fp64"]
    m = res["mixed"]
    if m["stabilisations_above_tolerance"] == 0:
        out["status"] = "ok"
        out["mixed"] = m
        out["speedup_vs_fp64"] = m["sweeps_per_s"] / res["fp64"]["sweeps_per_s"]
    else:
        # the chain's G left the bound: a rate for it would not be a rate of a valid chain
        out["status"] = "failed"
        out["mixed"] = {k: v for k, v in m.items() if k not in ("sweeps_per_s", "p50_ms", "p99_ms")}
        out["reason"] = (f"{m['stabilisations_above_tolerance']} of {m['stabilisations']} stabilisations "
                         f"exceeded max|dG| {tol:g} (largest {m['max_dG_per_stabilisation']:.3g})")
    return out


def cpu_all_cores(procs, sweeps=1):
    """The oracle on `procs` host processes at once, each an independent chain of the same
    workload (1 sweep each): aggregate sweeps/s of the host.  Child processes of this script
    (--cpu-child), which load the oracle only."""
    import subprocess
    t0 = time.perf_counter()
    ps = [subprocess.Popen([sys.executable, os.path.abspath(__file__), "--cpu-child", str(sweeps)],
                           stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True)
          for _ in range(procs)]
    outs = [p.communicate()[0] for p in ps]
    wall = time.perf_counter() - t0
    secs = [json.loads(o.strip().splitlines()[-1])["seconds"] for o in outs if o.strip()]
    if len(secs) != procs:
        return None
    return {"value": procs * sweeps / max(secs), "unit": "sweeps/s", "cores": procs,
            "kind": "port", "sample": f"{procs} processes x {sweeps} sweep, each its own chain; "
            f"slowest {max(secs):.1f} s, wall {wall:.1f} s"}


def cpu_baseline(pkg, model, par, field, sweeps):
    """The CPU oracle (FP64 restatement of the reference algorithm), one core."""
    sys.path.insert(0, ROOT)
    from oracle import oracle as O
    O.build()
    ch = O.OracleChain(len(model.l), field.flavors, par.slices, par.safe_mult, par.delta_tau,
                       field.alpha)
    ch.set_hopping(model.hopping_matrix())
    ch.set_field(field.conf)
    ch.set_seed(12345)
    ch.build_stack()
    t0 = time.perf_counter()
    ch.sweep(sweeps)
    dt = time.perf_counter() - t0
    return sweeps / dt, dt


def progress(msg):
    """One line on stderr as each leg ends (the JSON line on stdout comes at the very end)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-sweeps", type=int, default=2)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-mixed", action="store_true")
    ap.add_argument("--mixed-sweeps", type=int, default=3,
                    help="sweeps of the config-C FP64 and mixed-precision chains")
    ap.add_argument("--batch", type=int, default=64,
                    help="also time this many config-D chains swept together (0: skip)")
    ap.add_argument("--batch-sweeps", type=int, default=3, help="sweeps per chain of the batched leg")
    ap.add_argument("--serve", type=int, default=16,
                    help="also run this many config-E requests through the serving loop (0: skip)")
    ap.add_argument("--cpu-child", type=int, default=0, help=argparse.SUPPRESS)
    ap.add_argument("--cpu-procs", type=int, default=16,
                    help="host processes for the all-cores CPU baseline (0: skip)")
    ap.add_argument("--legs", default="headline,mixed,batched,serving,cpu",
                    help="comma-separated legs to run (headline always runs unless omitted here)")
    ap.add_argument("--profile-slices", type=int, default=0,
                    help="short run for hardware-counter profiles: build the stack, then this many "
                         "(sweep_spatial + propagate) slice steps, and exit")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one MI355X: independent chains shard across GPUs with no "
              "communication, so only --gpus 1 is run here", file=sys.stderr)
        sys.exit(2)

    pkg = load_pkg()
    model, par, field = config_b(pkg)
    n, flv, M = len(model.l), field.flavors, par.slices

    if args.cpu_child > 0:
        # one host process of the all-cores CPU baseline: the oracle only, no GPU
        cps, cdt = cpu_baseline(pkg, model, par, field, args.cpu_child)
        print(json.dumps({"sweeps": args.cpu_child, "seconds": cdt}))
        return

    if args.profile_slices > 0:
        ch = make_chain(pkg, model, par, field, "fp64")
        for _ in range(args.profile_slices):
            ch.sweep_spatial()
            ch.propagate()
        ch.synchronize()
        print(json.dumps({"profile_slices": args.profile_slices, "state": ch.state()}))
        ch.close()
        return

    legs = set(args.legs.split(","))
    if "headline" not in legs:
        # a single leg on its own (profiling runs): its JSON object alone
        out = {}
        if "batched" in legs:
            out["batched"] = batched_leg(pkg, args.batch, args.batch_sweeps)
        if "serving" in legs:
            out["serving"] = serve_leg(pkg, args.serve)
        if "mixed" in legs:
            out["mixed"] = mixed_leg(pkg, args.mixed_sweeps)
        print(json.dumps(out))
        return

    # ---- FP64 headline ------------------------------------------------------------------
    ch = make_chain(pkg, model, par, field, "fp64")
    total, lat, acc = timed_sweeps(ch, args.steps, args.warmup)
    sps = args.steps / total
    st = ch.stats()
    progress(f"headline: {sps:.2f} sweeps/s over {args.steps} sweeps")

    # ---- per-kernel device time (HIP events on the chain's stream) -------------------------
    ksteps = max(2, args.steps // 4)
    ch.set_kernel_timing(True)
    acc_k0 = ch.stats()["accepted"]
    t0 = time.perf_counter()
    ch.sweep(ksteps)
    t_events = time.perf_counter() - t0
    kt = ch.kernel_times()
    acc_k = ch.stats()["accepted"] - acc_k0
    ch.set_kernel_timing(False)
    kernels = {}
    win = 16 if n <= 256 else 32      # the library's default window (window_size, delay_depth 0)
    nwin = -(-n // win)
    launches = kt["local"][1]
    # one launch per slice (the default) or one per window
    wpl = nwin if launches and round(launches / ksteps) == 2 * par.slices else 1
    for kind, (ms, cnt) in kt.items():
        if cnt == 0:
            continue
        apl = acc_k / launches if kind in ("local", "local_apply") and launches else 0.0
        fl = kernel_work(kind, n, flv, win, apl, wpl if kind == "local" else 1)
        avg_s = ms / cnt / 1e3
        kernels[kind] = {"launches_per_sweep": cnt / ksteps, "avg_us": avg_s * 1e6,
                         "ms_per_sweep": ms / ksteps, "flop_per_launch": fl,
                         "tflops": fl / avg_s / 1e12 if avg_s > 0 else 0.0}
    dom = max(kernels, key=lambda k: kernels[k]["ms_per_sweep"])
    if dom == "local":
        # the window kernel is a latency chain of Metropolis decisions over a block held in
        # registers: its bound is memory (HBM bytes in and out per launch)
        nbytes = kernel_bytes("local", n, flv, win, wpl)
        kname = "window_slice_kernel" if wpl > 1 else "window_kernel"
        achieved = nbytes / (kernels[dom]["avg_us"] * 1e-6) / 1e9
        roofline = {"bound": "hbm", "kernel": kname, "achieved": achieved,
                    "peak": PEAK["hbm_gbs"], "unit": "GB/s", "frac": achieved / PEAK["hbm_gbs"],
                    "traffic": measured_traffic(kname),
                    "avg_launch_us": kernels[dom]["avg_us"], "bytes_per_launch": nbytes,
                    "windows_per_launch": wpl,
                    "traffic_source": "rocprofv3 FETCH_SIZE/WRITE_SIZE, profiles/dqmc_sweep_config_b_hbm_traffic.md"}
        # the HIP events around every launch add their own time: the committed rocprofv3
        # summary of the same workload gives the kernel's own duration
        rp = rocprof_avg_us(f"{kname}<double, {flv}, {win}>")
        if rp:
            roofline["avg_launch_us_rocprof"] = rp
            roofline["events_over_rocprof"] = kernels[dom]["avg_us"] / rp
            roofline["frac_rocprof"] = nbytes / (rp * 1e-6) / 1e9 / PEAK["hbm_gbs"]
        # the slice-product GEMM (two n x n items): measured HBM bytes against reading A and
        # B and writing C once per item
        gt = measured_traffic("gemm_kernel")
        if gt:
            alg = 3.0 * flv * n * n * 8
            roofline["gemm_traffic"] = {"kernel": "gemm_kernel<double, 1, 0, false, 32>",
                                        "traffic_bytes_per_launch": gt,
                                        "algorithmic_bytes_per_launch": alg, "ratio": gt / alg}
    else:
        peak = PEAK["fp32_mfma_tflops"] if dom == "gemm_f32" else PEAK["fp64_mfma_tflops"]
        achieved = kernels[dom]["tflops"]
        roofline = {"bound": "mfma", "kernel": dom, "achieved": achieved, "peak": peak,
                    "unit": "TFLOP/s", "frac": achieved / peak, "traffic": None,
                    "avg_launch_us": kernels[dom]["avg_us"],
                    "flop_per_launch": kernels[dom]["flop_per_launch"]}
    ch.close()
    progress(f"kernel times over {ksteps} sweeps")

    # ---- mixed precision at config C, beside the FP64 chain on the same field ---------------
    mixed = None
    if not args.skip_mixed and "mixed" in legs:
        mixed = mixed_leg(pkg, args.mixed_sweeps)
        progress(f"mixed: {mixed.get('status')}")

    # ---- batched form: config D, 64 chains in lockstep, one launch per kernel kind ---------
    batched = None
    if args.batch > 1 and "batched" in legs:
        batched = batched_leg(pkg, args.batch, args.batch_sweeps)
        batched["aggregate_vs_one_config_b_chain"] = batched["aggregate_sweeps_per_s"] / sps
        progress(f"batched: {batched['aggregate_sweeps_per_s']:.1f} sweeps/s")

    # ---- serving form: Poisson arrivals through the request loop -----------------------------
    serving = None
    if args.serve > 0 and "serving" in legs:
        serving = serve_leg(pkg, args.serve)
        progress(f"serving: {serving['wall_s']:.1f} s wall, {serving.get('status', 'complete')}")

    # ---- CPU baseline --------------------------------------------------------------------
    cpu = None
    if not args.skip_cpu and "cpu" in legs:
        cps, cdt = cpu_baseline(pkg, model, par, field, args.cpu_sweeps)
        cpu = {"value": cps, "unit": "sweeps/s", "cores": 1, "kind": "port",
               "sample": f"{args.cpu_sweeps} sweeps of the same chain (N=256, M=100, 2 flavours) "
                         f"after the stack build, {cdt:.1f} s, oracle/dqmc_oracle.c on one host core",
               "host_cores_available": os.cpu_count(),
               "gpu_over_cpu": sps / cps}
        progress(f"cpu oracle, one core: {cdt:.1f} s")
        if args.cpu_procs > 0:
            cpu["all_cores"] = cpu_all_cores(args.cpu_procs)
            progress("cpu oracle, all cores")

    # algorithmic work of a sweep: b n^3 (12 M + 46 K) + 2 b n^2 A (GEMM 2n^3, TRSM n^3,
    # UDT 3n^3; A accepted flips per sweep), against the FP64 matrix-core peak
    K = -(-M // par.safe_mult)
    f_sweep = flv * n ** 3 * (12 * M + 46 * K) + 2 * flv * n ** 2 * (acc / args.steps)
    sweep_roofline = {"flop_per_sweep": f_sweep, "achieved_tflops": f_sweep * sps / 1e12,
                      "peak_tflops": PEAK["fp64_mfma_tflops"],
                      "frac": f_sweep * sps / 1e12 / PEAK["fp64_mfma_tflops"]}

    out = {
        "metric": METRIC, "value": sps, "unit": "sweeps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": total / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "repulsive Hubbard 16x16 square, N=256, t=1, U=4, mu=0, beta=10, "
                               "dtau=0.1, M=100, magnetic Hirsch field (2 flavours), safe_mult=10, "
                               "1 chain, FP64",
                   "n_sites": n, "flavors": flv, "slices": M, "safe_mult": par.safe_mult},
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
        "acceptance": acc / (args.steps * 2 * n * M),
        "propagation_error_max": st["propagation_error_max"],
        "propagation_error_last": st["propagation_error_last"],
        "mixed": mixed,
        "batched": batched,
        "serving": serving,
        "roofline": roofline,
        "sweep_roofline": sweep_roofline,
        "kernels": kernels,
        "kernel_timing_region": {"sweeps": ksteps, "ms_per_sweep_with_events": t_events / ksteps * 1e3},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
    if _LEFT_RUNNING:
        # a server with requests past the deadline: closing it would wait for them
        sys.stdout.flush()
        sys.stderr.flush()
        os._exit(0)
