"""Continuous-batching request loop over the batched Davidson engine (jci_batch_*).

Requests (CI solves of different determinant counts) arrive at given times.  Between
lockstep iterations the loop admits every request that has arrived, as long as a slot
is free and its device memory fits the HBM budget; admitted requests share the GPU with
those already running (one sigma launch per iteration covers all of them) and leave as
soon as they converge, freeing slot and memory for the queue.  Latency of a request is
completion time minus arrival time (queueing included).

The synthetic request streams used by bench.py are generated here from a seed
(make_stream): instance sizes follow the batched configuration (orbitals n ~ U{8..14},
alpha = beta electrons ~ U{3..5}, roots ~ U{1..4}); arrival gaps are exponential.
"""
import math
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import (Batch, Context, Hamiltonian, Space, DIRECT_TABLES, GUESS_DIAG, PREC_FP64,
               RESTART_GDK, SYNTH_ENUC, synth_integrals)


def instance_spec(rng):
    n = int(rng.integers(8, 15))
    ne = int(rng.integers(3, 6))
    nr = int(rng.integers(1, 5))
    return dict(na=ne, nb=ne, norbs=n, ndocc=0, nactv=n, xlvl=0, rule=1, nroots=nr)


def make_stream(nreq, rate, seed, spec_fn=instance_spec):
    """Seeded Poisson stream: list of (arrival_s, spec, integral_seed)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(nreq):
        t += rng.exponential(1.0 / rate) if rate > 0 else 0.0
        out.append((t, spec_fn(rng), seed * 1000 + i))
    return out


def ndets_of(spec):
    from math import comb
    if spec["ndocc"] == 0 and spec["nactv"] == spec["norbs"]:
        return comb(spec["norbs"], spec["na"]) * comb(spec["norbs"], spec["nb"])
    return None


def solve_params(spec, N, restol=1e-6, nbuffer=1, etol=1e-9, kfac=8, buftol=1e-3,
                 restart=RESTART_GDK):
    """Davidson settings of a request (the same for the served and the one-at-a-time path):
    block of nroots + nbuffer tracked pairs, subspace up to kfac times that (at least 12);
    residual below restol and |r|^2 / gap below etol (eigenvalue error well under 1e-8 Ha
    also next to a close unrequested state).  The buffer pairs stop expanding once their
    residual is below buftol, and a full subspace restarts GD+k style (the lowest Ritz
    vectors and the previous iteration's ones): both cut the sigma applications per solve."""
    nr = spec["nroots"]
    nt = nr + nbuffer
    return dict(nroots=nr, restol=restol, krymin=min(nt + 1, N),
                krymax=min(max(kfac * nt, 12), N, 64), refdim=min(64, N), guess=GUESS_DIAG,
                totfrze=SYNTH_ENUC, nbuffer=nbuffer, etol=etol, buftol=buftol, restart=restart)


class Request:
    __slots__ = ("idx", "arrival", "spec", "seed", "space", "ham", "slot", "t_admit", "t_done",
                 "energies", "stats", "bytes", "params", "future")


def percentile(xs, q):
    xs = sorted(xs)
    if not xs:
        return None
    k = (len(xs) - 1) * q
    f = int(math.floor(k))
    c = min(f + 1, len(xs) - 1)
    return xs[f] + (xs[c] - xs[f]) * (k - f)


class Server:
    def __init__(self, ctx, capacity=64, budget_bytes=200e9, max_norbs=16, precision=PREC_FP64,
                 restol=1e-6, mixed_switch=1e-4, graphs=True, nbuffer=1, pool_bytes=24 << 30):
        self.ctx = ctx
        self.capacity = capacity
        # the sigma pool of the opposite-spin intermediates is part of the budget
        self.budget = budget_bytes - pool_bytes
        self.precision = precision
        self.restol = restol
        self.mixed_switch = mixed_switch
        self.nbuffer = nbuffer
        self.batch = Batch(ctx, capacity, max_norbs, graphs=graphs, pool_bytes=pool_bytes)
        # requests are built (integrals, strings, Hamiltonian, initial block) on a worker
        # thread with a context of its own, beside the steps of the running batch
        self.bctx = Context(ctx.device)
        self.builder = ThreadPoolExecutor(max_workers=1)

    def close(self):
        self.builder.shutdown(wait=True)
        self.batch.close()
        self.bctx.close()

    def _build(self, r):
        s = r.spec
        m1, m2 = synth_integrals(s["norbs"], r.seed)
        r.space = Space(self.bctx, s["na"], s["nb"], s["norbs"], s["ndocc"], s["nactv"],
                        s["xlvl"], s["rule"])
        r.ham = Hamiltonian(r.space, m1, m2, DIRECT_TABLES)
        r.params = self._params(r)
        r.ham.prepare_guess(r.params)

    def _params(self, req):
        sp = req.spec
        nr = sp["nroots"]
        N = req.space.ndets
        return req.ham.params(**solve_params(sp, N, self.restol, self.nbuffer),
                              precision=self.precision, mixed_switch=self.mixed_switch)

    def _estimate_bytes(self, spec):
        N = ndets_of(spec) or 10 ** 6
        krymax = solve_params(spec, N, self.restol, self.nbuffer)["krymax"]
        return Batch.bytes_for(N, krymax) + 64 * N

    def run(self, stream, time_scale=1.0):
        """Serve a stream of (arrival_s, spec, seed); arrivals are compressed by time_scale.
        Returns a list of finished Request objects (in completion order)."""
        reqs = []
        for i, (t, spec, seed) in enumerate(stream):
            r = Request()
            r.idx, r.arrival, r.spec, r.seed = i, t * time_scale, spec, seed
            r.space = r.ham = r.slot = None
            r.t_admit = r.t_done = None
            r.bytes = self._estimate_bytes(spec)
            reqs.append(r)
        waiting, building, active, done = [], [], {}, []
        used = 0.0
        nxt = 0
        t0 = time.perf_counter()
        while nxt < len(reqs) or waiting or building or active:
            now = time.perf_counter() - t0
            while nxt < len(reqs) and reqs[nxt].arrival <= now:
                waiting.append(reqs[nxt])
                nxt += 1
            # admission, first come first served within slots and HBM budget: the request is
            # built on the worker, then added to the batch here once its build has finished
            # (waiting for it only when nothing else could step)
            while (waiting and len(active) + len(building) < self.capacity
                   and used + waiting[0].bytes <= self.budget):
                r = waiting.pop(0)
                used += r.bytes
                r.future = self.builder.submit(self._build, r)
                building.append(r)
            while building and (building[0].future.done() or not active):
                r = building.pop(0)
                r.future.result()
                r.slot = self.batch.add(r.ham, r.params)
                r.t_admit = time.perf_counter() - t0
                active[r.slot] = r
            if not active:
                if waiting and not building:
                    # nothing runs and the head of the queue still does not fit
                    raise RuntimeError("request %d needs %.3g bytes, above the budget of %.3g"
                                       % (waiting[0].idx, waiting[0].bytes, self.budget))
                if nxt < len(reqs):
                    time.sleep(max(0.0, min(0.01, reqs[nxt].arrival - now)))
                continue
            self.batch.step()
            tnow = time.perf_counter() - t0
            for slot in self.batch.finished():
                if slot in active:
                    ev, _, st = self.batch.result(slot)
                    r = active.pop(slot)
                    r.t_done = tnow
                    r.energies = ev
                    r.stats = st
                    self.batch.release(slot)
                    r.ham.close()
                    r.space.close()
                    r.ham = r.space = None
                    used -= r.bytes
                    done.append(r)
        return done


def summarize(done, wall):
    lat = [r.t_done - r.arrival for r in done]
    return {"requests": len(done), "wall_s": wall, "solves_per_s": len(done) / wall if wall else None,
            "p50_s": percentile(lat, 0.5), "p99_s": percentile(lat, 0.99),
            "converged": sum(1 for r in done if r.stats["done"] == 1)}
