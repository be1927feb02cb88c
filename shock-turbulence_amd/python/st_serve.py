"""Ensembles and a serving loop over batches of same-size flow instances.

The reference runs one problem per MPI job.  Here many independent runs
share the GPU: instances of one grid size live in the slots of one batch
(st_batch_*), which advances all of them with one launch per pass and a
graph-replayed step, and the batches of different sizes run on their own
streams so their kernels overlap.

* Ensemble: a fixed set of instances advanced together for a number of
  steps (the batched-ensemble workload).
* Server: runs that arrive over time (a trace), are admitted into free
  slots as slots and memory free up, run to their own step counts and leave;
  time-to-completion is measured per run.

Both have a one-at-a-time reference on the plain path (one Solver per
run, st_march), which gives the same bits per run.  When the batches tune
their launch shapes (st_autotune), the plain path is given the same shapes
per grid size (st_pass_config_set): the lane layout sets the rounding.
"""
import collections
import time

import numpy as np

import st_amd

# device fields per slot (state, RK register, right-hand side, gradients,
# artificial-transport fields, partial sums, scratch): 57 x 8 bytes per point
BYTES_PER_POINT = 57 * 8


def slot_bytes(n):
    return BYTES_PER_POINT * n ** 3


class Run:
    """One flow: an n^3 periodic box from a seeded isotropic IC, `steps` steps."""

    def __init__(self, rid, n, steps, seed, k0=4.0, mt=0.3, re_lambda=40.0, arrival=0.0):
        self.rid, self.n, self.steps, self.seed = rid, n, steps, seed
        self.k0, self.mt, self.arrival = k0, mt, arrival
        self.Re = st_amd.re_for_taylor(re_lambda, k0, 1.0, mt)
        self.Q = None
        self.result = None
        self.admit = self.done = None
        self.time = None

    def make_ic(self):
        if self.Q is None:
            self.Q = st_amd.ic_isotropic(self.n, self.n, self.n, k0=self.k0, urms=1.0, mt=self.mt, seed=self.seed)
        return self.Q

    @property
    def work(self):
        """Grid points x RK stages."""
        return self.n ** 3 * 11 * self.steps


class Pool:
    """The batch of one grid size with its free-slot list."""

    def __init__(self, n, slots, params):
        self.n = n
        self.nslots = slots
        self.batch = st_amd.Batch(n, n, n, slots, **params)
        self.free = list(range(slots - 1, -1, -1))
        self.running = {}

    def tune(self, reps, run):
        """Launch shapes for this size, timed with run's state in every slot
        (the slots are free again afterwards)."""
        for slot in range(self.nslots):
            self.batch.load(slot, run.make_ic(), run.Re, 1)
        cfg = self.batch.autotune(reps)
        for slot in range(self.nslots):
            self.batch.release(slot)
        return cfg

    def admit(self, run, now):
        slot = self.free.pop()
        self.batch.load(slot, run.make_ic(), run.Re, run.steps)
        run.admit = now
        self.running[slot] = run

    def close(self):
        self.batch.close()


def _params(kw):
    p = dict(Pr=0.7, hyper_param=1, cfl_cal=0.5)
    p.update(kw)
    return p


def pool_sizes(runs, budget_bytes):
    """Slots per grid size: the concurrent demand of each size, scaled down
    together until the pools fit the memory budget."""
    want = collections.Counter(r.n for r in runs)
    scale = 1.0
    while True:
        slots = {n: max(1, int(c * scale)) for n, c in want.items()}
        if sum(slot_bytes(n) * s for n, s in slots.items()) <= budget_bytes or scale < 1e-3:
            return slots
        scale *= 0.9


class Ensemble:
    """Advance a fixed set of runs together: one pool per size.  With
    grouped (sizes of st_amd.GROUP_SIZES only), the pools step as one
    st_amd.Group, one launch per pass over all sizes, in the layouts of
    st_amd.narrow_configs; otherwise each pool steps on its own stream."""

    def __init__(self, runs, autotune=0, grouped=False, **kw):
        self.runs = runs
        self.pools = {}
        self.group = None
        for n, s in sorted(collections.Counter(r.n for r in runs).items()):
            self.pools[n] = Pool(n, s, _params(kw))
        for r in runs:
            self.pools[r.n].admit(r, 0.0)
        self.tuned = {}
        if grouped:
            if autotune:
                raise ValueError("a group takes fixed layouts: no autotune")
            self.group = st_amd.Group([p.batch for p in self.pools.values()])
            for n, p in self.pools.items():
                self.tuned[n] = st_amd.narrow_configs(p.batch)
        elif autotune:
            for n, p in self.pools.items():
                self.tuned[n] = p.batch.autotune(autotune)

    def march(self, nsteps):
        if self.group is not None:
            self.group.march(nsteps)
            self.group.sync()
            return
        for p in self.pools.values():
            p.batch.march(nsteps)
        for p in self.pools.values():
            p.batch.sync()

    def results(self):
        out = {}
        for p in self.pools.values():
            for slot, r in p.running.items():
                out[r.rid] = p.batch.store(slot)
        return out

    def close(self):
        if self.group is not None:
            self.group.close()
            self.group = None
        for p in self.pools.values():
            p.close()


class Server:
    """Continuous batching of arriving runs.

    Runs that have arrived are admitted first come, first served into free
    slots of their size's pool (the pools' slot counts come from the memory
    budget).  Each pool is driven on its own: whenever its stream has
    drained, the loop collects the runs that finished (a slot stops by
    itself at its run's step count), frees their slots, admits waiting runs
    and enqueues the next chunk of graph-replayed steps.  Chunks are sized
    to about `slice_s` seconds of that pool's work, so small and large
    pools keep their streams busy side by side and a finished run waits at
    most one chunk."""

    def __init__(self, runs, budget_bytes=200e9, slice_s=0.03, autotune=0, **kw):
        self.runs = sorted(runs, key=lambda r: r.arrival)
        self.slice_s = slice_s
        self.slots = pool_sizes(self.runs, budget_bytes)
        self.pools = {n: Pool(n, s, _params(kw)) for n, s in sorted(self.slots.items())}
        for r in self.runs:
            r.make_ic()                       # a run arrives with its initial state
        self.tuned = {}
        if autotune:                          # set-up, before the trace starts
            for n, p in self.pools.items():
                self.tuned[n] = p.tune(autotune, next(r for r in self.runs if r.n == n))

    def serve(self):
        pending = collections.deque(self.runs)
        queue = {n: collections.deque() for n in self.pools}
        left = len(self.runs)
        step_s = {n: None for n in self.pools}    # seconds per step, per pool (running estimate)
        issued = {n: None for n in self.pools}    # (time issued, steps) of the chunk in flight
        t0 = time.perf_counter()
        while left:
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now:
                r = pending.popleft()
                queue[r.n].append(r)
            progressed = False
            for n, p in self.pools.items():
                if issued[n] is not None and not p.batch.idle():
                    continue
                if issued[n] is not None:
                    t_iss, k = issued[n]
                    per = (time.perf_counter() - t0 - t_iss) / k
                    step_s[n] = per if step_s[n] is None else 0.5 * (step_s[n] + per)
                    issued[n] = None
                    step, t, _, _ = p.batch.status()
                    for slot, r in list(p.running.items()):
                        if step[slot] >= r.steps:
                            r.result = p.batch.store(slot)
                            r.time = t[slot]
                            r.done = time.perf_counter() - t0
                            p.batch.release(slot)
                            p.free.append(slot)
                            del p.running[slot]
                            left -= 1
                while queue[n] and p.free:
                    p.admit(queue[n].popleft(), time.perf_counter() - t0)
                if p.running:
                    k = 1 if step_s[n] is None else max(1, min(64, int(self.slice_s / max(step_s[n], 1e-6))))
                    k = min(k, max(r.steps for r in p.running.values()))
                    p.batch.march(k)
                    issued[n] = (time.perf_counter() - t0, k)
                    progressed = True
            if not progressed:
                if not any(issued.values()) and pending:
                    time.sleep(max(0.0, pending[0].arrival - (time.perf_counter() - t0)))
                else:
                    time.sleep(2e-4)
        return stats(self.runs, time.perf_counter() - t0)

    def close(self):
        for p in self.pools.values():
            p.close()


def stats(runs, makespan):
    lat = np.array([r.done - r.arrival for r in runs])
    work = sum(r.work for r in runs)
    return {
        "runs": len(runs),
        "makespan_s": makespan,
        "mpts_stage_per_s": work / makespan / 1e6,
        "latency_p50_s": float(np.percentile(lat, 50)),
        "latency_p99_s": float(np.percentile(lat, 99)),
        "latency_max_s": float(lat.max()),
    }


def plain_one_at_a_time(runs, configs=None, **kw):
    """The plain path: each run alone (st_march) in arrival order, first come
    first served; returns the same statistics and the per-run results.
    configs: {n: pass configurations} to apply per grid size."""
    results = {}
    service = {}
    for r in sorted(runs, key=lambda r: r.arrival):
        r.make_ic()
        s = st_amd.Solver(r.n, r.n, r.n, Re=r.Re, **_params(kw))   # set-up is not timed
        if configs and r.n in configs:
            s.set_pass_configs(configs[r.n])
        try:
            t1 = time.perf_counter()
            s.upload(r.Q)
            s.march(r.steps)
            results[r.rid] = s.download()
            service[r.rid] = time.perf_counter() - t1
        finally:
            s.close()
    total = sum(service.values())
    per_size = {}
    for r in runs:
        d = per_size.setdefault(str(r.n), {"runs": 0, "work": 0.0, "busy_s": 0.0})
        d["runs"] += 1
        d["work"] += r.work
        d["busy_s"] += service[r.rid]
    for d in per_size.values():
        d["mpts_stage_per_s"] = round(d["work"] / d["busy_s"] / 1e6, 1)
        d["work_share"] = round(d.pop("work") / sum(r.work for r in runs), 4)
        d["busy_s"] = round(d["busy_s"], 3)
    # first come, first served on one device
    clock = 0.0
    lat = []
    for r in sorted(runs, key=lambda r: r.arrival):
        clock = max(clock, r.arrival) + service[r.rid]
        lat.append(clock - r.arrival)
    lat = np.array(lat)
    work = sum(r.work for r in runs)
    st = {
        "runs": len(runs),
        "busy_s": total,
        "makespan_s": clock,
        "mpts_stage_per_s": work / total / 1e6,
        "latency_p50_s": float(np.percentile(lat, 50)),
        "latency_p99_s": float(np.percentile(lat, 99)),
        "latency_max_s": float(lat.max()),
        "per_size": per_size,
    }
    return st, results


def poisson_trace(nruns, rate, sizes=(48, 64, 96, 128), steps=(50, 400), seed=2024,
                  k0s=(3.0, 4.0, 6.0), mts=(0.2, 0.3, 0.5)):
    """Runs arriving as a Poisson process of the given rate (per second),
    grid size uniform over `sizes`, step count uniform in `steps`, and
    seeded IC parameters."""
    rng = np.random.default_rng(seed)
    t = 0.0
    runs = []
    for i in range(nruns):
        t += rng.exponential(1.0 / rate)
        runs.append(Run(i, int(rng.choice(sizes)), int(rng.integers(steps[0], steps[1] + 1)), seed=1000 + i,
                        k0=float(rng.choice(k0s)), mt=float(rng.choice(mts)), arrival=t))
    return runs


def ensemble_config3(nsteps):
    """48 instances: 16 x 48^3, 16 x 64^3, 12 x 96^3, 4 x 128^3, seeds 1..48,
    k0 and Mach number cycling over {3, 4, 6} and {0.2, 0.3, 0.5}."""
    sizes = [48] * 16 + [64] * 16 + [96] * 12 + [128] * 4
    k0s, mts = (3.0, 4.0, 6.0), (0.2, 0.3, 0.5)
    return [Run(i, n, nsteps, seed=i + 1, k0=k0s[i % 3], mt=mts[(i // 3) % 3]) for i, n in enumerate(sizes)]
