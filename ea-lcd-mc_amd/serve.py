"""Serving loop for EA tube simulations on one GPU: continuous batching.

Requests (seed, L, Lz, boundary condition, temperatures, sweeps) arrive
over time.  The server admits a request as soon as it has arrived and its
device footprint (ea_footprint) fits the HBM budget next to the running
ones; its lattice, couplings and generator states then stay on the device
until it completes.  All running requests advance together: the engines'
tubes, ordered by size, form one batch (ea_batch_run, one step of every tube
per graph replay) that runs in chunks of a few steps; between chunks,
finished requests leave (their memory is freed) and newly arrived ones join,
so requests that arrive while others are on the GPU overlap with them.  A
request's steps are the same whether it runs alone or in the mix: its
counter blocks, replica pairs and coins come from its own host streams, and
the chunks of a run are its consecutive steps.

The reference has no counterpart (it runs one tube size per process,
PROGRAM/Makefile:3-36); each request is that program's run of one job.
"""
import argparse
import json
import math
import sys
import threading
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

import eatube


@dataclass
class Request:
    rid: int
    seed: int
    L: int
    Lz: int
    bc: int
    nt: int
    sweeps: int
    arrival: float = 0.0  # seconds after the start of the run
    nr: int = 4
    ns: int = 1
    tmin: float = 0.6
    tmax: float = 1.6
    cluster: bool = True  # the reference program's step has the Houdayer
                          # move unless built with NO_CLUSTER (main.cu:182-184)

    def seeds(self):
        s = int(self.seed) * 4
        return (s + 1, s + 2, s + 3, s + 4)

    def attempts(self):
        return self.sweeps * self.L * self.L * self.Lz * self.nr * self.nt * self.ns


@dataclass
class Result:
    rid: int
    arrival: float
    admitted: float
    finished: float
    spins: np.ndarray = None
    perms: np.ndarray = None
    esum: np.ndarray = None
    energies: np.ndarray = None

    @property
    def latency(self):
        return self.finished - self.arrival


def poisson_requests(n, rate, seed=1, sizes=(8, 16, 24, 32), nts=(16, 24, 32),
                     sweeps=(1000, 100000), lz_max=16, cluster=True):
    """n requests at Poisson arrivals of `rate` per second: L from `sizes`,
    Lz a multiple of 8 in [L, lz_max*L], PBC or OBC, NT from `nts`, sweeps
    log-uniform in [sweeps[0], sweeps[1]]; each step has the Houdayer move
    when `cluster`."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        L = int(rng.choice(sizes))
        Lz = 8 * int(rng.integers(L // 8, lz_max * L // 8 + 1))
        sw = int(round(math.exp(rng.uniform(math.log(sweeps[0]),
                                            math.log(sweeps[1])))))
        out.append(Request(rid=i, seed=1 + 7919 * i, L=L, Lz=Lz,
                           bc=int(rng.integers(0, 2)), nt=int(rng.choice(nts)),
                           sweeps=sw, arrival=t, cluster=bool(cluster)))
    return out


class _Luts:
    """LUTs by (nt, tmin, tmax): built once (quad precision on the host)."""

    def __init__(self):
        self.cache = {}
        self.lock = threading.Lock()

    def get(self, r):
        key = (r.nt, r.tmin, r.tmax)
        with self.lock:
            if key not in self.cache:
                betas = eatube.betas_linear(r.tmin, r.tmax, r.nt)
                self.cache[key] = (betas, eatube.make_luts(betas))
            return self.cache[key]


def prepare(r, luts):
    """The host half of request r's start (eatube.Prepared): no GPU work,
    so it runs on a worker thread while the GPU serves the running ones."""
    luts.get(r)
    return eatube.Prepared(r.L, r.Lz, r.nr, r.nt, r.ns, bc=r.bc, seeds=r.seeds())


def start(r, luts, prepared=None):
    """Device-resident state of request r, initialised as the reference
    program does (eatube.Simulation)."""
    betas, lut = luts.get(r)
    return eatube.Simulation(r.L, r.Lz, r.nr, r.nt, r.ns, bc=r.bc, betas=betas,
                             lut=lut, seeds=r.seeds(), job_id=r.rid % (1 << 14),
                             prepared=prepared)


def footprint(r, luts):
    betas, _ = luts.get(r)
    return eatube.footprint(r.L, r.Lz, r.nr, r.nt, r.ns, r.bc, betas)


class Server:
    """Continuous batching of requests within `budget` bytes of HBM.

    chunk      steps every running request advances between admissions
    max_tubes  most requests on the GPU at once
    keep       keep each result's final state (spins, perms, energy sums)
    workers    threads preparing arrived requests on the host (prepare_fn)
               while the GPU serves; 0: each request is prepared inside
               start_fn at its admission
    shape      launch shape of the batches (eatube.Batch.set_shape
               arguments); the default suits requests of one sample: one
               thread per position makes its random words and updates the
               sample (sweep kind 1)
    progress   called as progress(seconds, finished, running, pending) at
               most every progress_every seconds, between chunks
    """

    def __init__(self, budget, chunk=32, max_tubes=256, keep=False,
                 start_fn=start, batch_fn=eatube.Batch, footprint_fn=footprint,
                 clock=time.perf_counter, workers=0, prepare_fn=prepare,
                 shape=(64, 512, 1, 0), progress=None, progress_every=30.0):
        self.budget = int(budget)
        self.chunk = int(chunk)
        self.max_tubes = int(max_tubes)
        self.keep = keep
        self.start_fn, self.batch_fn = start_fn, batch_fn
        self.footprint_fn, self.clock = footprint_fn, clock
        self.workers, self.prepare_fn = int(workers), prepare_fn
        self.shape = tuple(shape) if shape else None
        self.progress, self.progress_every = progress, float(progress_every)
        self.luts = _Luts()
        self.stats = dict(chunks=0, batches_built=0, max_running=0,
                          max_bytes=0, deferred=0, t_start=0.0, t_batch=0.0,
                          t_schedule=0.0, t_run=0.0, t_finish=0.0)

    def run(self, requests):
        """Serve `requests` (arrival times relative to now); returns the
        results in request order.  A request whose footprint exceeds the
        budget is refused (ValueError) before anything starts; if serving
        fails later, the running engines, the batch and the worker pool are
        released before the error propagates."""
        need_of = {}
        for r in requests:
            need_of[r.rid] = self.footprint_fn(r, self.luts)
            if need_of[r.rid] > self.budget:
                raise ValueError("request %d needs %d bytes, more than the "
                                 "budget" % (r.rid, need_of[r.rid]))
        pending = sorted(requests, key=lambda r: (r.arrival, r.rid))
        pool = ThreadPoolExecutor(self.workers) if self.workers > 0 else None
        running = {}  # rid -> [request, sim, steps left, bytes, admitted]
        batch = None
        try:
            results, batch = self._serve(pending, need_of, pool, running)
        finally:
            if batch is not None:
                batch.close()
            for ent in running.values():
                ent[1].engine.close()
            if pool is not None:
                pool.shutdown(wait=True, cancel_futures=True)
        return [results[r.rid] for r in sorted(requests, key=lambda r: r.rid)]

    def _serve(self, pending, need_of, pool, running):
        futures = {}  # rid -> preparation in flight
        t0 = self.clock()
        used = 0
        results = {}
        batch, members = None, ()
        told = 0.0
        try:
            while pending or running:
                now = self.clock() - t0
                if self.progress is not None and now - told >= self.progress_every:
                    told = now
                    self.progress(now, len(results), len(running), len(pending))
                if pool is not None:  # arrived requests start preparing at once
                    for r in pending:
                        if r.arrival > now:
                            break
                        if r.rid not in futures:
                            futures[r.rid] = pool.submit(self.prepare_fn, r,
                                                         self.luts)
                # admission: arrived (and prepared) requests, in arrival
                # order, while they fit
                k = 0
                while k < len(pending) and pending[k].arrival <= now and \
                        len(running) < self.max_tubes:
                    r = pending[k]
                    if pool is not None and not futures[r.rid].done():
                        k += 1
                        continue
                    need = need_of[r.rid]
                    if used + need > self.budget:
                        self.stats["deferred"] += 1
                        break
                    pending.pop(k)
                    ts = self.clock()
                    if pool is None:
                        sim = self.start_fn(r, self.luts)
                    else:
                        sim = self.start_fn(r, self.luts,
                                            futures.pop(r.rid).result())
                    running[r.rid] = [r, sim, r.sweeps, need, self.clock() - t0]
                    self.stats["t_start"] += self.clock() - ts
                    used += need
                self.stats["max_running"] = max(self.stats["max_running"],
                                                len(running))
                self.stats["max_bytes"] = max(self.stats["max_bytes"], used)
                if not running:
                    if pending and pending[0].arrival > self.clock() - t0:
                        time.sleep(max(0.0, pending[0].arrival -
                                       (self.clock() - t0)))
                    else:
                        time.sleep(0.0005)  # arrived requests still preparing
                    continue
                # the running requests, ordered by size, are the batch's tubes
                order = tuple(sorted(running, key=lambda k: (
                    running[k][0].L * running[k][0].L * running[k][0].Lz, k)))
                ts = self.clock()
                if order != members:
                    if batch is not None:
                        batch.close()
                        batch = None
                    batch = self.batch_fn([running[k][1].engine for k in order])
                    if self.shape and hasattr(batch, "set_shape"):
                        batch.set_shape(*self.shape)
                    members = order
                    self.stats["batches_built"] += 1
                tb = self.clock()
                scheds = []
                for k in order:
                    r, sim, left = running[k][:3]
                    scheds.append(sim.schedule(min(self.chunk, left),
                                               cluster=r.cluster))
                tc = self.clock()
                batch.run(scheds)
                self.stats["chunks"] += 1
                done = self.clock() - t0
                self.stats["t_batch"] += tb - ts
                self.stats["t_schedule"] += tc - tb
                self.stats["t_run"] += self.clock() - tc
                ts = self.clock()
                for k, sch in zip(order, scheds):
                    running[k][2] -= sch.n_it
                    if running[k][2] == 0:
                        r, sim, _, need, adm = running[k]
                        res = Result(r.rid, r.arrival, adm, done)
                        if self.keep:
                            es, _ = sim.engine.read_measures()
                            res.spins = sim.spins()
                            res.perms = sim.engine.download_perms()
                            res.esum = es
                            res.energies = sim.engine.energy()
                        del running[k]
                        sim.engine.close()
                        used -= need
                        results[r.rid] = res
                self.stats["t_finish"] += self.clock() - ts
        except BaseException:
            if batch is not None:
                batch.close()
            raise
        return results, batch


def summary(requests, results, wall):
    lat = np.array([x.latency for x in results])
    work = sum(r.attempts() for r in requests)
    return {"requests": len(results),
            "p50_latency_s": float(np.percentile(lat, 50)),
            "p99_latency_s": float(np.percentile(lat, 99)),
            "mean_latency_s": float(lat.mean()),
            "wall_s": wall,
            "attempts_per_ns": work / (wall * 1e9)}


def record(r, x):
    """One finished request as a JSON-able record."""
    return {"rid": r.rid, "L": r.L, "Lz": r.Lz, "bc": r.bc, "nt": r.nt,
            "sweeps": r.sweeps, "attempts": r.attempts(), "arrival": x.arrival,
            "admitted": x.admitted, "finished": x.finished, "latency": x.latency}


def merge(paths):
    """The summary of several runs' records files (segments of one stream):
    latency percentiles over all requests, throughput = all attempts over
    the sum of the segments' walls."""
    recs, wall = [], 0.0
    for p in paths:
        with open(p) as f:
            d = json.load(f)
        recs += d["records"]
        wall += d["wall_s"]
    lat = np.array([x["latency"] for x in recs])
    return {"requests": len(recs), "segments": len(paths),
            "p50_latency_s": float(np.percentile(lat, 50)),
            "p99_latency_s": float(np.percentile(lat, 99)),
            "mean_latency_s": float(lat.mean()), "wall_s": wall,
            "attempts": float(sum(x["attempts"] for x in recs)),
            "attempts_per_ns": sum(x["attempts"] for x in recs) / (wall * 1e9)}


def main(argv=None):
    """Serve a Poisson stream of requests (or a segment of one: requests
    first .. first+count-1 of the same seeded stream, their arrival times
    taken from the previous request's), print progress and a JSON summary,
    and write every request's record to --out; --merge prints the summary
    of several such files."""
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--requests", type=int, default=2000)
    ap.add_argument("--rate", type=float, default=0.5, help="arrivals per second")
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--first", type=int, default=0)
    ap.add_argument("--count", type=int, default=0, help="0: to the end")
    ap.add_argument("--sweeps", default="1000,100000", help="log-uniform range")
    ap.add_argument("--chunk", type=int, default=128)
    ap.add_argument("--budget-gb", type=float, default=64.0)
    ap.add_argument("--workers", type=int, default=8)
    ap.add_argument("--out", default="")
    ap.add_argument("--merge", nargs="*", default=None)
    args = ap.parse_args(argv)
    if args.merge is not None:
        print(json.dumps(merge(args.merge)))
        return 0
    lo, hi = (int(x) for x in args.sweeps.split(","))
    reqs = poisson_requests(args.requests, args.rate, seed=args.seed, sweeps=(lo, hi))
    last = args.requests if args.count <= 0 else min(args.requests, args.first + args.count)
    seg = reqs[args.first:last]
    t_prev = reqs[args.first - 1].arrival if args.first > 0 else 0.0
    for r in seg:
        r.arrival -= t_prev

    def say(t, done, running, pending):
        print("%.0f s: %d finished, %d running, %d waiting" % (t, done, running, pending),
              flush=True)

    srv = Server(budget=int(args.budget_gb * 2 ** 30), chunk=args.chunk,
                 workers=args.workers, progress=say, progress_every=60.0)
    t0 = time.perf_counter()
    res = srv.run(seg)
    wall = time.perf_counter() - t0
    out = summary(seg, res, wall)
    out.update({"first": args.first, "count": len(seg), "rate_per_s": args.rate,
                "seed": args.seed, "sweeps_range": [lo, hi], "chunk_steps": args.chunk,
                "max_running": srv.stats["max_running"],
                "batches_built": srv.stats["batches_built"]})
    if args.out:
        with open(args.out, "w") as f:
            json.dump({"summary": out, "wall_s": wall,
                       "records": [record(r, x) for r, x in zip(seg, res)]}, f)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
