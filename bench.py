#!/usr/bin/env python3
"""Throughput of the EA tube Monte Carlo engine on one MI355X.

A step is one pass of the reference's inner loop over all samples (main.cu
:149-228 with ptfr = houfr = mesfr = 1): a Metropolis sweep of every
(site, replica, temperature, sample), the energy, parallel tempering, the
Houdayer move (when the workload has it) and the per-step energy measure.
The metric is the reference's own unit (timing.c:27-34):

    spin-flip attempts/ns = steps * V * NR * NT * S / wall_ns

Default workload: the OBC tube L=16, Lz=64, 4 replicas, 24 temperatures in
[0.6, 1.6], S=31 samples, Houdayer at the lround(24/2)=12 lowest
temperatures every step.  Inputs (spins, couplings, PT state) are resident
on the GPU before the timed region; the host-side counter blocks and replica
pairs of every step are generated inside it.

Prints one JSON line.  --gpus above 1 is refused: this build measures one
GPU.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "ea-lcd-mc_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

import eatube  # noqa: E402

HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # the configuration the headline metric is quoted on
    "obc_L16_Lz64": dict(L=16, Lz=64, nr=4, nt=24, tmin=0.6, tmax=1.6,
                         bc=eatube.OBC, cluster=True),
    "plain_L8_Lz16": dict(L=8, Lz=16, nr=2, nt=16, tmin=0.7, tmax=1.6,
                          bc=eatube.PBC, cluster=False),
    "large_L48_Lz384": dict(L=48, Lz=384, nr=4, nt=32, tmin=0.5, tmax=1.6,
                            bc=eatube.PBC, cluster=True),
}
SEEDS = (20240601, 20240602, 20240603, 20240604)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def b_alg_sweep(S, nr, nt):
    """Algorithmic HBM bytes per spin-flip attempt of the sweep: own word
    read+write, the opposite parity read once, 3 random words shared by the
    S samples, 6 coupling words shared by NR*NT words (per 64 attempts)."""
    return 0.375 + 0.375 / S + 0.75 / (nr * nt)


def kernel_bytes(kind, w, S, V, ctm, sweep_halves=1):
    """algorithmic bytes of one launch of each kernel kind (a sweep launch
    makes one half-sweep, or both with sweep kind 3)"""
    nr, nt = w["nr"], w["nt"]
    if kind == "sweep":
        return b_alg_sweep(S, nr, nt) * V * nr * nt * S / 2.0 * sweep_halves
    if kind == "houdayer":
        # q read and both replicas written: 0.5 B per labelled site when the
        # labels stay on chip; 8.5 B when they round-trip HBM
        per_site = 0.5 if V <= 32768 else 8.5
        return per_site * V * ctm * S * 2
    if kind == "energy":
        return V * nr * nt * S / 8.0 + 6 * V / 8.0 * S  # spins + couplings
    if kind == "pt":
        return S * nr * (2080 * 2 + nt * 4 + nt * 2 + (nt - 1) * 24)
    return 0.0


def copy_calibration(mib):
    """A device-to-device copy of `mib` MiB (the runtime's blit kernel),
    whose HBM traffic is known: mib MiB read and mib MiB written.  Profiled
    beside the engine's kernels it converts FETCH_SIZE / WRITE_SIZE counter
    readings into bytes."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    hip.hipMemset(a, 1, C.c_size_t(n))
    for _ in range(3):
        assert hip.hipMemcpy(b, a, C.c_size_t(n), 3) == 0  # device to device
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)


def measured_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary
    profiles/traffic_<workload>.txt (two rocprofv3 --pmc runs, FETCH_SIZE
    and WRITE_SIZE, scaled by the copy calibration), or None."""
    path = os.path.join(ROOT, "profiles", "traffic_%s.txt" % workload)
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) >= 3 and f[0] == "traffic_bytes_per_launch" and f[1] == kernel:
            return float(f[2])
    return None


def parity_check(w):
    """Two steps of two samples of the workload against the oracle."""
    import orc
    betas = eatube.betas_linear(w["tmin"], w["tmax"], w["nt"])
    args = (w["L"], w["Lz"], w["nr"], w["nt"], 2)
    e = eatube.Simulation(*args, bc=w["bc"], betas=betas, seeds=SEEDS)
    o = orc.Sim(*args, bc=w["bc"], betas=betas, seeds=SEEDS)
    e.steps(2, cluster=w["cluster"])
    o.steps(2, cluster=w["cluster"])
    ok = np.array_equal(e.spins(), o.spins()) and np.array_equal(
        e.engine.download_perms().ravel(), o.field("wctb"))
    o.energy()
    ok = ok and np.array_equal(e.engine.energy().ravel(), o.field("ener"))
    e.engine.close()
    return ok


def mixed_tubes(n=48, seed=7):
    """The uneven mix: n tubes, L from {8, 16, 24} (multiples of 8 only:
    MSC needs L/4 even), Lz from {L, 4L, 8L}, half PBC and half OBC, 4
    replicas, 24 temperatures in [0.6, 1.6], per-tube seeds."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        L = int(rng.choice([8, 16, 24]))
        Lz = L * int(rng.choice([1, 4, 8]))
        out.append((L, Lz, i % 2))
    return out


def batched_compare(steps, samples=1, ntubes=48):
    """The same tubes from the same seeds, advanced `steps` steps (sweep +
    PT + Houdayer + measure) tube by tube with ea_run, and as one batch
    (ea_batch_run); states must end bit-identical."""
    tubes = mixed_tubes(ntubes)
    nt, nr = 24, 4
    betas = eatube.betas_linear(0.6, 1.6, nt)
    lut = eatube.make_luts(betas)

    def sims():
        return [eatube.Simulation(L, Lz, nr, nt, samples, bc=bc, betas=betas,
                                  lut=lut, seeds=(9000 + 4 * k, 9001 + 4 * k,
                                                  9002 + 4 * k, 9003 + 4 * k),
                                  job_id=k)
                for k, (L, Lz, bc) in enumerate(tubes)]
    plain, twins = sims(), sims()
    attempts = steps * sum(L * L * Lz for L, Lz, _ in tubes) * nr * nt * samples
    # plain form: one tube after the other, kernels launched per tube
    for p in plain:  # warm-up call (module load, plans)
        p.steps(1, cluster=True)
        p.engine.sync()
    scheds = [p.schedule(steps, cluster=True) for p in plain]
    t0 = time.perf_counter()
    for p, sch in zip(plain, scheds):  # each tube's run completes before the next
        p.engine.run(sch)               # starts (ea_run returns once enqueued, and
        p.engine.sync()                 # engines have streams of their own)
    t_plain = time.perf_counter() - t0
    # batched form, in the launch shape the autotuner finds fastest for
    # this mix (timed on scratch tubes of the same configurations)
    shape, shape_times = eatube.tune_shape([t.engine for t in twins], steps=10)
    batch = eatube.Batch([t.engine for t in twins])
    batch.set_shape(*shape)
    batch.run([t.schedule(1, cluster=True) for t in twins])
    scheds = [t.schedule(steps, cluster=True) for t in twins]
    t0 = time.perf_counter()
    batch.run(scheds)
    t_batch = time.perf_counter() - t0
    same = True
    for p, t in zip(plain, twins):
        same = same and np.array_equal(p.spins(), t.spins()) and \
            np.array_equal(p.engine.download_perms(), t.engine.download_perms())
    batch.close()
    for x in plain + twins:
        x.engine.close()
    return {"tubes": len(tubes), "steps": steps, "samples_per_tube": samples,
            "sites": int(sum(L * L * Lz for L, Lz, _ in tubes)),
            "plain_value": attempts / (t_plain * 1e9),
            "batched_value": attempts / (t_batch * 1e9),
            "speedup": t_plain / t_batch, "bit_exact": bool(same),
            "launch_shape": list(shape),
            "tuner_ms_per_step": {"/".join(map(str, k)): 1e3 * v / 10
                                  for k, v in shape_times.items()},
            "unit": "spin-flip attempts/ns"}


def serving(nreq, rate, sweeps, budget_gb=64.0, chunk=128):
    """Poisson stream of simulation requests through the continuous-batching
    server: p50/p99 completion latency (arrival to completion) and
    whole-run throughput."""
    import serve
    reqs = serve.poisson_requests(nreq, rate, seed=5, sweeps=sweeps)
    workers = min(8, int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1)
    srv = serve.Server(budget=int(budget_gb * 2 ** 30), chunk=chunk, workers=workers)
    t0 = time.perf_counter()
    res = srv.run(reqs)
    wall = time.perf_counter() - t0
    out = serve.summary(reqs, res, wall)
    out.update({"rate_per_s": rate, "sweeps_range": list(sweeps),
                "chunk_steps": chunk, "hbm_budget_gb": budget_gb,
                "prepare_workers": workers,
                "max_running": srv.stats["max_running"],
                "batches_built": srv.stats["batches_built"],
                "chunks": srv.stats["chunks"],
                "host_seconds": {k[2:]: round(v, 3) for k, v in srv.stats.items()
                                 if k.startswith("t_")},
                "request_shape": "L in {8,16,24,32}, Lz multiple of 8 in "
                                 "[L,16L], PBC|OBC, NT in {16,24,32} in "
                                 "[0.6,1.6], 4 replicas, 1 sample, sweeps "
                                 "log-uniform in [%d,%d], sweep+PT+Houdayer per "
                                 "step" % sweeps})
    return out


def cpu_baseline(w, S, budget_s):
    """The oracle (scalar C restatement of the reference algorithm) on the
    host: all the job's cores (OpenMP over samples, chains and cluster
    instances, OMP_NUM_THREADS or every core) and one core, each over a
    bounded number of steps of the same workload (half the budget each)."""
    import orc
    betas = eatube.betas_linear(w["tmin"], w["tmax"], w["nt"])
    V = w["L"] * w["L"] * w["Lz"]
    cores = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1

    def rate(threads, seconds):
        orc.set_threads(threads)
        o = orc.Sim(w["L"], w["Lz"], w["nr"], w["nt"], S, bc=w["bc"],
                    betas=betas, seeds=SEEDS)
        t0 = time.perf_counter()
        o.steps(1, cluster=w["cluster"])
        t1 = time.perf_counter() - t0
        n = int(max(1, min(2000, seconds / max(t1, 1e-6))))
        t0 = time.perf_counter()
        o.steps(n, cluster=w["cluster"])
        dt = time.perf_counter() - t0
        return n * V * w["nr"] * w["nt"] * S / (dt * 1e9), n, dt

    v_all, n_all, dt_all = rate(cores, budget_s / 2)
    v_one, n_one, dt_one = rate(1, budget_s / 2)
    orc.set_threads(1)
    return {"value": v_all, "unit": "spin-flip attempts/ns", "cores": cores,
            "kind": "port", "one_core_value": v_one,
            "sample": "%d steps of the same workload (all %d samples) in "
                      "%.1f s on %d threads; %d steps in %.1f s on one core"
                      % (n_all, S, dt_all, cores, n_one, dt_one)}


def measure(name, S, steps, warmup, timing_steps, tune=True, shape=None, tune_steps=50):
    """One workload at S samples: the launch shape (tuned, or given as
    "P,C,K,X"), `warmup` untimed steps, `steps` timed steps of the engine
    with its inputs resident (the host counter blocks and replica pairs are
    made inside the timed region: the first step's, then the others' while
    the first runs), then per-kernel durations from HIP events over
    `timing_steps` steps and each kernel's algorithmic GB/s."""
    w = WORKLOADS[name]
    V = w["L"] * w["L"] * w["Lz"]
    betas = eatube.betas_linear(w["tmin"], w["tmax"], w["nt"])
    sim = eatube.Simulation(w["L"], w["Lz"], w["nr"], w["nt"], S, bc=w["bc"],
                            betas=betas, seeds=SEEDS)
    eng = sim.engine
    ctm = eng.ctm if w["cluster"] else 0
    shp, shape_times = eatube.DEFAULT_SHAPE, {}
    if shape:
        shp = tuple(int(x) for x in shape.split(","))
    elif tune:
        shp, shape_times = eatube.tune_shape([eng], cluster=w["cluster"], steps=tune_steps)
        log(name, "launch shape", shp, {k: round(v * 1e3 / tune_steps, 4)
                                         for k, v in shape_times.items()})
    eng.set_shape(*shp)
    # untimed warmup in the pattern of the timed region below (a one-step
    # call, then one call for the rest), so that its code paths are warm
    log(name, "warmup", warmup)
    if warmup > 0:
        eng.run(sim.schedule(1, cluster=w["cluster"]))
    if warmup > 1:
        eng.run(sim.schedule(warmup - 1, cluster=w["cluster"]))
    eng.sync()
    log(name, "timed", steps)
    t0 = time.perf_counter()
    # the first step's inputs go first, so that the GPU starts while the host
    # draws the other steps' counter blocks and pairs (same streams, same
    # order: the state is the one a single schedule of all steps gives)
    first = sim.schedule(1, cluster=w["cluster"])
    eng.run(first)
    if steps > 1:
        rest = sim.schedule(steps - 1, cluster=w["cluster"])
        eng.run(rest)
    eng.sync()
    wall = time.perf_counter() - t0
    value = steps * V * w["nr"] * w["nt"] * S / (wall * 1e9)
    # per-kernel durations, HIP events on the engine's stream
    eng.set_timing(True)
    sim.steps(timing_steps, cluster=w["cluster"])
    tm = eng.read_timing()
    eng.set_timing(False)
    kern = {}
    halves = 2 if shp[2] == 3 else 1  # half-sweeps per sweep launch
    for k, (ms, n) in tm.items():
        if n:
            b = kernel_bytes(k, w, S, V, ctm, halves)
            kern[k] = {"avg_us": 1e3 * ms / n, "launches": n,
                       "share": 0.0, "GBs": b / (1e-3 * ms / n) / 1e9,
                       "frac": b / (1e-3 * ms / n) / 1e9 / HBM_PEAK_GBS}
    tot = sum(v["avg_us"] * v["launches"] for v in kern.values())
    for v in kern.values():
        v["share"] = v["avg_us"] * v["launches"] / tot
    dom = max(kern, key=lambda k: kern[k]["share"])
    ach = kern[dom]["GBs"]
    roof = {"bound": "hbm", "kernel": dom, "achieved": ach,
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
            "traffic": measured_traffic(name, dom),
            "bytes_per_launch": kernel_bytes(dom, w, S, V, ctm, halves),
            "avg_launch_us": kern[dom]["avg_us"]}
    eng.close()
    return {"workload": name, "value": value, "unit": "spin-flip attempts/ns",
            "steps": steps, "warmup": warmup, "ms_per_step": 1e3 * wall / steps,
            "wall_s": wall, "samples": S, "houdayer_betas": ctm,
            "shape": shp, "shape_times": shape_times, "kernels": kern, "roofline": roof}


def sub_record(name, S, steps, warmup, timing_steps):
    """A further configuration measured like the headline, in brief."""
    r = measure(name, S, steps, warmup, timing_steps, tune_steps=max(2, min(50, steps)))
    w = WORKLOADS[name]
    return {"value": r["value"], "unit": r["unit"], "ms_per_step": r["ms_per_step"],
            "steps": steps, "warmup": warmup,
            "config": {"L": w["L"], "Lz": w["Lz"], "bc": "OBC" if w["bc"] else "PBC",
                       "replicas": w["nr"], "temperatures": w["nt"],
                       "T_range": [w["tmin"], w["tmax"]], "samples": S,
                       "houdayer": bool(w["cluster"]), "houdayer_betas": r["houdayer_betas"]},
            "launch_shape": list(r["shape"]),
            "kernels": {k: {"avg_us": v["avg_us"], "GBs": v["GBs"], "frac": v["frac"],
                            "share": v["share"]} for k, v in r["kernels"].items()},
            "roofline": r["roofline"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--workload", default="obc_L16_Lz64",
                    choices=sorted(WORKLOADS))
    ap.add_argument("--samples", type=int, default=31)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-parity", action="store_true")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default launch shape")
    ap.add_argument("--shape", default=None,
                    help="launch shape P,C,K,X (sweep positions, cluster "
                         "threads, sweep kind, XCD order) instead of tuning")
    ap.add_argument("--timing-steps", type=int, default=50)
    ap.add_argument("--batch-steps", type=int, default=200,
                    help="steps of the 48-tube mixed batch comparison (0: skip)")
    ap.add_argument("--serve-requests", type=int, default=24,
                    help="requests of the serving run (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=0.5,
                    help="Poisson arrivals per second of the serving run")
    ap.add_argument("--serve-chunk", type=int, default=128,
                    help="steps every running request advances between admissions")
    ap.add_argument("--no-sub", action="store_true",
                    help="skip the plain and large configurations' sub-records")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="also run a MIB-MiB device copy (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU; runs on %d GPUs are not part of "
              "this build" % max(args.gpus, int(os.environ.get("WORLD_SIZE",
                                                               "1"))),
              file=sys.stderr)
        sys.exit(2)
    if eatube.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
    w = WORKLOADS[args.workload]
    S = args.samples

    parity = None
    if not args.no_parity:
        parity = parity_check(w)
        log("parity vs oracle:", parity)
        if not parity:
            print(json.dumps({"metric": "spin-flip attempts/ns",
                              "error": "bit-exact check against the oracle "
                                       "failed; no timing reported"}))
            sys.exit(1)

    head = measure(args.workload, S, args.steps, args.warmup, args.timing_steps,
                   tune=not args.no_tune, shape=args.shape)
    shape, shape_times, kern, roof = (head["shape"], head["shape_times"], head["kernels"],
                                      head["roofline"])
    value, wall, ctm = head["value"], head["wall_s"], head["houdayer_betas"]
    sweep = kern.get("sweep")
    out = {
        "metric": "spin-flip attempts/ns",
        "value": value,
        "unit": "spin-flip attempts/ns",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (seeded couplings and spins, reference seeding chain)",
        "config": {"workload": args.workload, "L": w["L"], "Lz": w["Lz"],
                   "bc": "OBC" if w["bc"] else "PBC", "replicas": w["nr"],
                   "temperatures": w["nt"], "T_range": [w["tmin"], w["tmax"]],
                   "samples": S, "houdayer": bool(w["cluster"]),
                   "houdayer_betas": ctm if w["cluster"] else 0,
                   "step": "sweep+energy+PT" + ("+Houdayer" if w["cluster"]
                                                else "") + "+measure"},
        "parity_vs_oracle": parity,
        "launch_shape": {"sweep_positions": shape[0],
                         "cluster_threads": shape[1],
                         "sweep_kind": shape[2], "sweep_xcd": shape[3],
                         "tuner_ms_per_step": {"/".join(map(str, k)): 1e3 * v / 50
                                               for k, v in shape_times.items()}},
        "kernels": kern,
        "roofline": roof,
    }
    if sweep:
        out["sweep_roofline"] = {"achieved": sweep["GBs"], "peak": HBM_PEAK_GBS,
                                 "frac": sweep["GBs"] / HBM_PEAK_GBS,
                                 "avg_launch_us": sweep["avg_us"]}
    if not args.no_sub:
        # the plain small tube (no cluster moves) and the large tube
        log("sub-record plain_L8_Lz16")
        out["plain_L8_Lz16"] = sub_record("plain_L8_Lz16", S, 2000, 100, 50)
        log("sub-record large_L48_Lz384")
        out["large_L48_Lz384"] = sub_record("large_L48_Lz384", S, 20, 3, 3)
    if args.batch_steps > 0:
        log("batched vs plain, 48 mixed tubes")
        out["batched"] = batched_compare(args.batch_steps)
    if args.serve_requests > 0:
        log("serving %d requests" % args.serve_requests)
        out["serving"] = serving(args.serve_requests, args.serve_rate,
                                 (1000, 100000), chunk=args.serve_chunk)
    if not args.no_cpu:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(w, S, args.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
