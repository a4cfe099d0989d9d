"""Continuous-batching request loop over the batched Davidson engine (jci_batch_*).

Requests (CI solves of different determinant counts) arrive at given times.  Between
lockstep iterations the loop admits every request that has arrived, as long as a slot
is free and its device memory fits the HBM budget; admitted requests share the GPU with
those already running (one sigma launch per iteration covers all of them) and leave as
soon as they converge, freeing slot and memory for the queue.  Latency of a request is
completion time minus arrival time (queueing included).

Requests the batch engine does not take (more orbitals than the batch was sized for: the
truncated MRCI-SD spaces of 16 electrons in 60 orbitals) run on a lane of their own: a worker
thread with a context (stream) of its own builds the space and runs the whole Davidson
solve, one such request at a time, while the batch keeps stepping beside it; admission
counts its device memory against the same budget.

The synthetic request streams used by bench.py are generated here from a seed
(make_stream): instance sizes follow the batched configuration (orbitals n ~ U{8..14},
alpha = beta electrons ~ U{3..5}, roots ~ U{1..4}); arrival gaps are exponential.
make_stream(..., mrci_share=0.1) replaces that share of the requests by MRCI-SD 16e/60o
solves with 1-3 roots.
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


def mrci_spec(rng, nvirt=48):
    """MRCI-SD 16 electrons in 12 + nvirt orbitals: 4 doubly occupied + CAS(8e,8o) + nvirt
    external, excitation level 2, MRCI-SD rule (603,651,972 determinants at 48 external),
    roots ~ U{1..3}."""
    return dict(na=8, nb=8, norbs=12 + nvirt, ndocc=4, nactv=8, xlvl=2, rule=1,
                nroots=int(rng.integers(1, 4)))


def make_stream(nreq, rate, seed, spec_fn=instance_spec, mrci_share=0.0, mrci_nvirt=48):
    """Seeded Poisson stream: list of (arrival_s, spec, integral_seed).  With mrci_share > 0
    round(mrci_share * nreq) requests, at least one, chosen by their own generator (so the
    full-CI requests are those of the stream without the share), are MRCI-SD solves of 16
    electrons in 12 + mrci_nvirt orbitals (mrci_spec)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(nreq):
        t += rng.exponential(1.0 / rate) if rate > 0 else 0.0
        out.append((t, spec_fn(rng), seed * 1000 + i))
    if mrci_share > 0 and nreq > 0:
        mr = np.random.default_rng(seed + 7)
        nm = max(1, int(round(mrci_share * nreq)))
        for i in sorted(mr.choice(nreq, size=nm, replace=False)):
            out[i] = (out[i][0], mrci_spec(mr, mrci_nvirt), out[i][2])
    return out


def ndets_of(spec):
    from math import comb
    if spec["ndocc"] == 0 and spec["nactv"] == spec["norbs"]:
        return comb(spec["norbs"], spec["na"]) * comb(spec["norbs"], spec["nb"])
    return None


def lane_vector_bytes(N, krymax, kb):
    """Device bytes of a lane solve over N determinants: basis and sigma vectors (2 krymax),
    new directions and the determinant-major sigma input (2 kb), the block-transposed input
    (kb, FP64) and its FP32 copy (kb), the diagonal, and the excitation tables (bounded by
    120 bytes per determinant for excitation level 2: 104 at 16e/60o)."""
    return N * (8 * (2 * krymax + 2 * kb) + 12 * kb + 8 + 120)


def lane_solve_params(spec, N, restol=1e-6, etol=1e-9, buftol=1e-3, budget=None, kfac=4):
    """Davidson settings of a lane request (a space far larger than the batched ones): one
    buffer pair for a single root, none above; a subspace of kfac times the tracked pairs
    (at least 8), or the largest that keeps lane_vector_bytes within `budget`, but not below
    twice the tracked pairs (at least 6); GD+k restarts, diagonal guess.  With three tracked
    pairs the subspace of 12 converges in 71 iterations where 9 takes 99 and 6 stalls
    (MRCI-SD 16e/20o, 3 roots), so memory goes to the subspace first."""
    nr = spec["nroots"]
    nbuf = 1 if nr == 1 else 0
    nt = nr + nbuf
    kmax, kmin = max(kfac * nt, 8), max(2 * nt, 6)
    while budget is not None and kmax > kmin and lane_vector_bytes(N, kmax, nt) > budget:
        kmax -= 1
    return dict(nroots=nr, restol=restol, krymin=min(nt + 1, N), krymax=min(kmax, N),
                refdim=min(64, N), guess=GUESS_DIAG, totfrze=SYNTH_ENUC, nbuffer=nbuf,
                etol=etol, buftol=buftol, restart=RESTART_GDK)


def lane_bytes(spec, N, restol=1e-6, budget=None):
    """Device bytes a lane request holds with lane_solve_params(..., budget)."""
    p = lane_solve_params(spec, N, restol, budget=budget)
    return lane_vector_bytes(N, p["krymax"], p["nroots"] + p["nbuffer"])


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
        self.max_norbs = max_norbs
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
        # the lane of requests the batch does not take (see the module docstring)
        self.lctx = Context(ctx.device)
        self.lane = ThreadPoolExecutor(max_workers=1)
        self.t0 = None

    def close(self):
        self.builder.shutdown(wait=True)
        self.lane.shutdown(wait=True)
        self.batch.close()
        self.bctx.close()
        self.lctx.close()

    def tune_tiles(self, spec, seed, reps=5):
        """Autotune the batch's dense-string tile widths on an instance of spec (the largest
        full-CI request class of the stream); returns (bn64, bn32)."""
        m1, m2 = synth_integrals(spec["norbs"], seed)
        sp = Space(self.ctx, spec["na"], spec["nb"], spec["norbs"], spec["ndocc"],
                   spec["nactv"], spec["xlvl"], spec["rule"])
        h = Hamiltonian(sp, m1, m2, DIRECT_TABLES)
        try:
            return self.batch.tune_tiles(h, nvec=1, reps=reps)
        finally:
            h.close()
            sp.close()

    def _on_lane(self, spec):
        return spec["norbs"] > self.max_norbs

    def _solve_lane(self, r):
        """The whole solve of a lane request, on the lane's context."""
        s = r.spec
        m1, m2 = synth_integrals(s["norbs"], r.seed)
        r.space = Space(self.lctx, s["na"], s["nb"], s["norbs"], s["ndocc"], s["nactv"],
                        s["xlvl"], s["rule"])
        r.ham = Hamiltonian(r.space, m1, m2, DIRECT_TABLES)
        r.t_admit = time.perf_counter() - self.t0
        ev, _, st = r.ham.solve(precision=self.precision, mixed_switch=self.mixed_switch,
                                vectors=False, check=False,
                                **lane_solve_params(s, r.space.ndets, self.restol,
                                                    budget=self.budget))
        self.lctx.sync()
        r.t_done = time.perf_counter() - self.t0
        r.energies = ev
        r.stats = dict(st, done=1 if st["converged"] else 2)
        r.ham.close()
        r.space.close()
        r.ham = r.space = None

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
        if self._on_lane(spec):
            sp = Space(self.ctx, spec["na"], spec["nb"], spec["norbs"], spec["ndocc"],
                       spec["nactv"], spec["xlvl"], spec["rule"])
            N = sp.ndets
            sp.close()
            return lane_bytes(spec, N, self.restol, budget=self.budget)
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
        lane = None                     # the lane request running, if any
        used = 0.0
        nxt = 0
        t0 = self.t0 = time.perf_counter()
        while nxt < len(reqs) or waiting or building or active or lane:
            now = time.perf_counter() - t0
            while nxt < len(reqs) and reqs[nxt].arrival <= now:
                waiting.append(reqs[nxt])
                nxt += 1
            # admission (admit): a batch request is built on the worker, then added to the
            # batch here once its build has finished (waiting for it only when nothing else
            # could step); a lane request starts on the lane when the lane is free
            lane_pick, picks, used = admit(waiting, lambda q: self._on_lane(q.spec),
                                           lane is not None, used, self.budget,
                                           self.capacity - len(active) - len(building))
            if lane_pick is not None:
                lane = waiting[lane_pick]
                lane.future = self.lane.submit(self._solve_lane, lane)
            for i in picks:
                r = waiting[i]
                r.future = self.builder.submit(self._build, r)
                building.append(r)
            if lane_pick is not None or picks:
                taken = set(picks) | ({lane_pick} if lane_pick is not None else set())
                waiting = [r for j, r in enumerate(waiting) if j not in taken]
            if lane is not None and lane.future.done():
                lane.future.result()
                used -= lane.bytes
                done.append(lane)
                lane = None
                continue
            while building and (building[0].future.done() or not active):
                r = building.pop(0)
                r.future.result()
                r.slot = self.batch.add(r.ham, r.params)
                r.t_admit = time.perf_counter() - t0
                active[r.slot] = r
            if not active:
                if waiting and not building and lane is None:
                    # nothing runs and the head of the queue still does not fit
                    raise RuntimeError("request %d needs %.3g bytes, above the budget of %.3g"
                                       % (waiting[0].idx, waiting[0].bytes, self.budget))
                if lane is not None:
                    time.sleep(0.005)
                elif nxt < len(reqs):
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


def admit(waiting, on_lane, lane_busy, used, budget, slots_free):
    """One admission pass over the waiting queue (objects with .bytes), first come first
    served within the HBM budget and the batch's free slots: returns (index of the request
    that starts on the lane or None, indices of the requests for the batch in queue order,
    bytes in use after them).  A lane request waiting for the busy lane does not hold up
    the batch requests behind it (each kind keeps its order); the first request that does
    not fit the budget stops the pass, and so does a batch request without a free slot."""
    lane_pick, picks = None, []
    for i, r in enumerate(waiting):
        lane_req = on_lane(r)
        if lane_req and (lane_busy or lane_pick is not None):
            continue
        if used + r.bytes > budget:
            break
        if lane_req:
            lane_pick = i
            used += r.bytes
            continue
        if slots_free <= 0:
            break
        picks.append(i)
        used += r.bytes
        slots_free -= 1
    return lane_pick, picks, used


def summarize(done, wall):
    lat = [r.t_done - r.arrival for r in done]
    return {"requests": len(done), "wall_s": wall, "solves_per_s": len(done) / wall if wall else None,
            "p50_s": percentile(lat, 0.5), "p99_s": percentile(lat, 0.99),
            "converged": sum(1 for r in done if r.stats["done"] == 1)}
