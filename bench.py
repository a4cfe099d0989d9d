"""Benchmark: Davidson solves/s and p99 latency at 1e-8 Ha, mixed vs plain FP64, one GPU.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload served|batched|fci12]

Workloads (synthetic integrals from seeds, resident on the GPU before each solve starts):
  served   (default) a seeded Poisson stream (mean 200 requests/s, seed 400) of K CI
           solves drawn like the batched configuration: full CI, orbitals n ~ U{8..14},
           alpha = beta electrons ~ U{3..5}, roots ~ U{1..4}, residual 1e-6.  Timed on the
           same stream for
           (a) the served path: continuous batching (one sigma launch per lockstep
               iteration for every resident request), mixed precision (FP32 sigma ->
               FP64 refinement), admission by slots and a 200 GB HBM budget  -> value;
           (b) the plain FP64 GPU path: one request at a time in arrival order;
           (c) the served path in FP64 (isolates batching from precision).
           A step is one request; ms_per_step = wall time of (a) / K.
  batched  the same distribution, all K instances present at t = 0 (seed 300).
  fci12    full CI 4a4b / 12 orbitals (245,025 determinants), 1 root: K solves per path.
  mrci     MRCI-SD 16e/60o (603,651,972 determinants), --roots R, --precision fp64 | mixed |
           both: one solve per precision, per-root |dE| of mixed against FP64.

Each root of every path is compared with a tight FP64 GPU solve of the same instance
(residual 1e-9, FP64 sigma); max |dE| is reported and the target is <= 1e-8 Ha.  The
roofline is that of the dominant kernel, the dense-string FP64 sigma, on the largest request
(MFMA-bound; FP32 beside it), timed with HIP events on the library's stream; the pair-list
sigma's HBM roofline on the full-CI 4a4b/12 space is reported beside it.  The CPU baseline
is the oracle's OpenMP sigma timed for one pass over full CI 4a4b/12 on this machine's host
cores and converted to solves/s through the workload's total sigma work; the reference's own
compute_hv, timed against the same pass where the reference tree is present
(oracle/time_reference.py -> profiles/cpu_reference_fci12.json), gives the estimate beside it.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM = 8.0e12        # MI355X HBM3E, bytes/s
METRIC = "Davidson solves/s + p99 latency at 1e-8 Ha, mixed vs plain FP64, 1 GPU"
METHOD_NAMES = {1: "pairs64", 2: "pairs32", 3: "pairs_sym64", 4: "direct64", 5: "direct32",
                6: "fast64", 7: "fast32", 8: "fci64", 9: "fci32", 10: "gemm64", 11: "gemm32"}


def log(msg):
    print("bench: " + msg, file=sys.stderr, flush=True)


def percentile(xs, q):
    xs = sorted(xs)
    if not xs:
        return None
    k = (len(xs) - 1) * q
    f = int(k)
    c = min(f + 1, len(xs) - 1)
    return xs[f] + (xs[c] - xs[f]) * (k - f)


LANE_NORBS = 16     # requests with more orbitals run on the serving loop's lane


def solve_kw(J, spec, N, restol):
    """The request's Davidson settings, the same on every path: the batched ones for the
    full-CI requests, the lane ones for the MRCI-SD requests."""
    from jayci_amd.serve import solve_params, lane_solve_params
    kw = (lane_solve_params(spec, N, restol) if spec["norbs"] > LANE_NORBS
          else solve_params(spec, N, restol))
    kw["vectors"] = False
    return kw


def make_instance(J, ctx, spec, seed, flags):
    m1, m2 = J.synth_integrals(spec["norbs"], seed)
    sp = J.Space(ctx, spec["na"], spec["nb"], spec["norbs"], spec["ndocc"], spec["nactv"],
                 spec["xlvl"], spec["rule"])
    return sp, J.Hamiltonian(sp, m1, m2, flags)


class c_stdout_to_stderr:
    """Send what C code writes to file descriptor 1 to stderr while active (the solver's
    per-iteration lines); stdout carries only the JSON line."""

    def __enter__(self):
        sys.stdout.flush()
        self.saved = os.dup(1)
        os.dup2(2, 1)
        return self

    def __exit__(self, *exc):
        import ctypes
        ctypes.CDLL(None).fflush(None)
        os.dup2(self.saved, 1)
        os.close(self.saved)
        return False


def fci_pairs(spec):
    """(N, unique off-diagonal pairs) of a full-CI space: c = 2a(n-a) + 2C(a,2)C(n-a,2)
    + (a(n-a))^2 connections per row for a = alpha = beta electrons."""
    from math import comb
    n, a, b = spec["norbs"], spec["na"], spec["nb"]
    N = comb(n, a) * comb(n, b)
    c = (a * (n - a) + b * (n - b) + comb(a, 2) * comb(n - a, 2) + comb(b, 2) * comb(n - b, 2)
         + a * (n - a) * b * (n - b))
    return N, N * c // 2


def cpu_baseline(J, work):
    """CPU baseline timed on this machine's host cores: the oracle's OpenMP sigma
    (oracle/jayci_oracle.c, the CPU restatement of the sigma rules of
    mpi_source/pdavidson.c:1964-2067 with mpi_source/action_util.c:167-258) for one pass over
    full CI 4a4b/12 (245,025 determinants, seed 100).  Its rate in Hamiltonian elements per
    second (P_u + N per pass) converts the workload's sigma work into solves/s:
    work = (solves, elements), elements = sum over the plain FP64 GPU solves of (P_u + N) x
    sigma applications.  Reference code never runs here; the reference's own compute_hv
    (oracle/_ref) was timed against the same oracle pass in the build container, and that
    ratio (profiles/cpu_reference_fci12.json) gives the reference estimate beside it."""
    import numpy as np
    import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    m1, m2 = J.synth_integrals(12, 100)
    osp = oracle.Space(4, 4, 12)
    v = np.random.default_rng(0).standard_normal(osp.ndets)
    osp.sigma(m1, m2, v, nthreads=threads)            # warm (page-in, thread pool)
    ts = []
    for _ in range(3):
        t0 = time.time()
        osp.sigma(m1, m2, v, nthreads=threads)
        ts.append(time.time() - t0)
    t = min(ts)
    N, P_u = fci_pairs({"norbs": 12, "na": 4, "nb": 4})
    rate = (P_u + N) / t
    solves, elements = work
    out = {"value": solves * rate / elements, "unit": "solves/s", "cores": threads,
           "kind": "port",
           "sample": "one sigma pass of the oracle's OpenMP restatement (gcc -O2 -fopenmp, %d "
                     "threads) over full CI 4a4b/12 (N=%d, P_u=%d): %.3f s = %.3g matrix "
                     "elements/s; the workload's %d plain solves need %.4g elements (P_u + N "
                     "per sigma application)" % (threads, N, P_u, t, rate, solves, elements)}
    cal = os.path.join(ROOT, "profiles", "cpu_reference_fci12.json")
    if os.path.exists(cal):
        c = json.load(open(cal))
        ratio = c["sigma_ratio_reference_over_oracle"]
        out["reference_estimate"] = {
            "value": out["value"] / ratio, "unit": "solves/s",
            "calibration": "reference compute_hv / oracle sigma = %.2f on the same pass, "
                           "both %d threads on %s; reference generate_cimap %.0f s and one "
                           "compute_hv pass %.1f s on that instance (sigma-only conversion, "
                           "the map excluded)" % (ratio, c["threads"], c["cpu"],
                                                  c["generate_cimap_s"], c["compute_hv_pass_s"])}
    return out


def roofline_of(J, ctx, spec, seed):
    """FP64 pair-list sigma (k_spmv<double>, one vector) on one instance: algorithmic bytes
    B = P_u*(8+4) + N*2*8 + 8*N over the HIP-event time of one launch.  On a full-CI
    instance the dense-string FP64 sigma of the same vector is timed too and reported beside
    B / peak, the shortest time a sigma that streams the pair list could take."""
    sp, h = make_instance(J, ctx, spec, seed, J.PAIRS_FP64 | J.DIRECT_TABLES)
    N = sp.ndets
    P_u = h.npairs // 2
    B = P_u * 12 + N * 16 + 8 * N
    ms = h.bench_sigma(J.SIGMA_PAIRS64, 1, 20)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "sigma_pairs64_fci12_traffic.json")
    if os.path.exists(tpath):
        t = json.load(open(tpath))
        if t.get("workload") == "fci12" and N == 245025:
            traffic = t.get("bytes_per_launch")
    ach = B / (ms * 1e-3)
    # the same sigma by the dense-string path (full CI only): it moves none of the pair
    # bytes, so it is set against the time the pair-list roofline allows for B
    dense_ms = h.bench_sigma(J.SIGMA_GEMM64, 1, 20) if spec["ndocc"] == 0 and \
        spec["nactv"] == spec["norbs"] else None
    out = {"bound": "hbm", "achieved": ach / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
           "frac": ach / PEAK_HBM, "traffic": traffic,
           "kernel": "k_spmv<double> (FP64 pair list, 1 vector), N=%d (%da%db/%d)"
                     % (N, spec["na"], spec["nb"], spec["norbs"]),
           "ms_per_launch": ms, "algorithmic_bytes": B, "pairs_unique": P_u,
           "roofline_ms_at_peak": B / PEAK_HBM * 1e3,
           "dense_string_fp64_ms": dense_ms}
    h.close()
    sp.close()
    return out


PEAK_MFMA = {64: 78.6, 32: 157.3}   # MI355X dense MFMA peaks, TFLOP/s


def roofline_dense(J, ctx, spec, seed):
    """Dense-string sigma (k_gE + k_gss + k_gga, one vector) on one full-CI instance, FP64
    and FP32: algorithmic FLOPs 4*P_u + 2*N (two FMAs per unique connected pair, one per
    diagonal element) over the HIP-event time of one sigma.  Beside it, the FLOPs the GEMMs
    issue, 2*La*m1*N (opposite-spin product: La = a(n-a) + 1 rows per alpha string against
    all m1 pairs of (P|Q)) + 2*(Na^2*Nb + Na*Nb^2) (dense same-spin products), which count
    the structural zeros of those products."""
    from math import comb
    sp, h = make_instance(J, ctx, spec, seed, J.DIRECT_TABLES)
    n = spec["norbs"]
    na, nb = comb(n, spec["na"]), comb(n, spec["nb"])
    N = na * nb
    m1 = n * (n + 1) // 2
    La = spec["na"] * (n - spec["na"]) + 1
    gemm_flops = 2.0 * La * m1 * N + 2.0 * (na * na * nb + na * nb * nb)
    # algorithm-independent count: two FMAs per unique connected pair and one per diagonal
    # element, 4 P_u + 2 N per vector
    _, P_u = fci_pairs(spec)
    flops = 4.0 * P_u + 2.0 * N
    out = {}
    traffic = {}
    tpath = os.path.join(ROOT, "profiles", "sigma_dense_5a5b14_traffic.json")
    if os.path.exists(tpath):
        t = json.load(open(tpath))
        if t.get("ndets") == N:
            traffic = {64: t.get("bytes_per_launch_fp64"), 32: t.get("bytes_per_launch_fp32")}
    for prec, meth in ((64, J.SIGMA_GEMM64), (32, J.SIGMA_GEMM32)):
        ms = h.bench_sigma(meth, 1, 20)
        ach = flops / (ms * 1e-3) / 1e12
        out[prec] = {"bound": "mfma", "achieved": ach, "peak": PEAK_MFMA[prec],
                     "unit": "TFLOP/s", "frac": ach / PEAK_MFMA[prec],
                     "traffic": traffic.get(prec),
                     "kernel": "dense-string sigma %s (k_gE + k_gss + k_gga), N=%d (%da%db/%d)"
                               % ("FP64" if prec == 64 else "FP32", N, spec["na"], spec["nb"], n),
                     "ms_per_launch": ms, "algorithmic_flops": flops,
                     "flop_count": "4*P_u + 2*N (P_u unique connected pairs)",
                     "gemm_flops": gemm_flops,
                     "gemm_frac": gemm_flops / (ms * 1e-3) / 1e12 / PEAK_MFMA[prec]}
    h.close()
    sp.close()
    return out


def roofline_table(J, ctx, spec, seed):
    """Matrix-free table sigma (k_block_transpose, k_sigma_fast and k_sigma_fast_beta
    launches of one pass, default settings), FP64 and FP32, one vector, on one MRCI-SD
    instance of the stream.  HBM-bound by its governing definition: the compulsory bytes of
    a pass that cannot materialise its pair list, N*(2*k*b + 8) (vector read, sigma
    written, diagonal read; b bytes per vector element), over the HIP-event time of one
    pass; the connected entries per second (N times the mean connections of 64 sampled
    rows) beside it.  `traffic` is the rocprof byte count of the same pass."""
    import numpy as np
    sp, h = make_instance(J, ctx, spec, seed, J.DIRECT_TABLES)
    N = sp.ndets
    rows = np.random.default_rng(0).integers(0, N, size=64)
    cbar = float(np.mean([len(c) for c, _, _, _ in h.rows(rows)]))
    traffic = {}
    tpath = os.path.join(ROOT, "profiles", "sigma_tables_mrcisd_16e20o_traffic.json")
    if os.path.exists(tpath):
        t = json.load(open(tpath))
        if t.get("ndets") == N:
            traffic = {64: t.get("bytes_per_launch_fp64")}
    out = {}
    for prec, meth, b in ((64, J.SIGMA_FAST64, 8), (32, J.SIGMA_FAST32, 4)):
        ms = h.bench_sigma(meth, 1, 5)
        byt = N * (2 * b + 8)
        ach = byt / (ms * 1e-3) / 1e9
        out[prec] = {"bound": "hbm", "achieved": ach, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                     "frac": ach * 1e9 / PEAK_HBM, "traffic": traffic.get(prec),
                     "kernel": "table sigma %s (k_block_transpose + k_sigma_fast + "
                               "k_sigma_fast_beta), N=%d (MRCI-SD 16e/%do)"
                               % ("FP64" if prec == 64 else "FP32", N, spec["norbs"]),
                     "ms_per_launch": ms, "algorithmic_bytes": byt,
                     "byte_count": "N*(2*k*b + 8): vector read, sigma written, diagonal read",
                     "connected_entries_per_s": N * cbar / (ms * 1e-3)}
    h.close()
    sp.close()
    return out


def run_plain(J, ctx, stream, restol, precision):
    """One request at a time in arrival order (no batching).  Returns (results, wall)."""
    out = []
    t0 = time.perf_counter()
    for i, (arr, s, seed) in enumerate(stream):
        now = time.perf_counter() - t0
        if arr > now:
            time.sleep(arr - now)
        sp, h = make_instance(J, ctx, s, seed, J.DIRECT_TABLES)
        ev, _, st = h.solve(precision=precision, **solve_kw(J, s, sp.ndets, restol))
        h.close()
        sp.close()
        out.append((i, time.perf_counter() - t0 - arr, list(ev), st))
    return out, time.perf_counter() - t0


def reference_energies(J, ctx, stream):
    """Tight FP64 GPU solves (residual 1e-8, FP64 sigma, two buffer roots converged like the
    requested ones, subspace of 8*(roots + 2) collapsing onto Ritz vectors, up to 1000
    iterations) of every instance; the energy error of such a solve is ~|r|^2 / gap, far below
    1e-8 Ha.  Returns (energies, largest final residual norm)."""
    ref, rmax = [], 0.0
    for arr, s, seed in stream:
        sp, h = make_instance(J, ctx, s, seed, J.DIRECT_TABLES)
        kw = solve_kw(J, s, sp.ndets, 1e-8)
        if s["norbs"] > LANE_NORBS:
            # MRCI-SD requests: the lane settings at residual 1e-8 with one buffer root and a
            # subspace of six times the tracked pairs
            nt = s["nroots"] + 1
            kw.update(nbuffer=1, krymin=min(nt + 1, sp.ndets),
                      krymax=min(max(6 * nt, 8), sp.ndets), maxiter=1000, buftol=0.0)
        else:
            kw.update(nbuffer=2, krymin=min(s["nroots"] + 3, sp.ndets),
                      krymax=min(max(8 * (s["nroots"] + 2), 24), sp.ndets, 64), maxiter=1000,
                      buftol=0.0, restart=J.RESTART_RITZ)
        ev, _, st = h.solve(precision=J.PREC_FP64, check=False, **kw)
        ref.append(list(ev))
        rmax = max(rmax, max(st["rnorm"]))
        h.close()
        sp.close()
    return ref, rmax


def bench_stream(J, ctx, args, rate, seed):
    from jayci_amd import serve
    share = args.mrci_share if rate else 0.0
    stream = serve.make_stream(args.steps, rate, seed, mrci_share=share,
                               mrci_nvirt=args.mrci_nvirt)
    warm = serve.make_stream(args.warmup, 0.0, seed + 1) if args.warmup else []
    ndets = []
    for _, s, _ in stream:
        n = serve.ndets_of(s)
        if n is None:
            sp = J.Space(ctx, s["na"], s["nb"], s["norbs"], s["ndocc"], s["nactv"], s["xlvl"],
                         s["rule"])
            n = sp.ndets
            sp.close()
        ndets.append(n)
    nmrci = sum(1 for _, s, _ in stream if s["norbs"] > LANE_NORBS)
    log("stream: %d requests (%d MRCI-SD), %d determinants in total, largest %d"
        % (len(stream), nmrci, sum(ndets), max(ndets)))

    phases = {}
    tiles = {}
    # roofline instance and tile-tuning instance: the largest full-CI request
    big = max((i for i in range(len(stream)) if stream[i][1]["norbs"] <= LANE_NORBS),
              key=lambda i: ndets[i])

    def served(precision, st):
        # the opposite-spin pool shrinks to 8 GiB when MRCI-SD requests share the budget
        srv = serve.Server(ctx, capacity=256, budget_bytes=200e9, max_norbs=LANE_NORBS,
                           precision=precision, restol=1e-6,
                           pool_bytes=(8 << 30) if nmrci else (24 << 30))
        try:
            tiles[precision] = srv.tune_tiles(stream[big][1], stream[big][2])
            ctx.sync()
            t0 = time.perf_counter()
            done = srv.run(st)
            ctx.sync()
            wall = time.perf_counter() - t0
            sig = srv.batch.sigma_seconds
            phases[precision] = {k: round(v, 4) for k, v in srv.batch.phase_seconds.items()}
        finally:
            srv.close()
        return done, wall, sig

    if warm:
        served(J.PREC_MIXED, warm)
        run_plain(J, ctx, warm, 1e-6, J.PREC_FP64)
    done_m, wall_m, sig_m = served(J.PREC_MIXED, stream)
    log("served mixed: %.3f s (sigma %.3f s)" % (wall_m, sig_m))
    plain, wall_p = run_plain(J, ctx, stream, 1e-6, J.PREC_FP64)
    log("plain fp64 one at a time: %.3f s" % wall_p)
    done_b, wall_b, sig_b = served(J.PREC_FP64, stream)
    log("served fp64: %.3f s" % wall_b)
    ref, ref_r = reference_energies(J, ctx, stream)
    log("reference energies done (largest residual %.2e)" % ref_r)

    plain_ev = {i: ev for i, _, ev, _ in plain}
    for i in range(len(stream)):
        if ref[i] is None:
            ref[i] = plain_ev[i]

    def err_of(idx, ev):
        return max(abs(ev[k] - ref[idx][k]) for k in range(len(ev)))

    def worst(items):
        """Three largest errors: (request, n, electrons, roots, |dE|, final residuals)."""
        out = sorted(items, key=lambda x: -x[1])[:3]
        return [{"request": i, "norbs": stream[i][1]["norbs"], "nelec_a": stream[i][1]["na"],
                 "roots": stream[i][1]["nroots"], "abs_err_Ha": e,
                 "rnorm": [float("%.3g" % r) for r in rn]} for i, e, rn in out]
    err_m = max(err_of(r.idx, r.energies) for r in done_m)
    err_b = max(err_of(r.idx, r.energies) for r in done_b)
    err_p = max(err_of(i, ev) for i, _, ev, _ in plain)
    worst_m = worst([(r.idx, err_of(r.idx, r.energies), r.stats["rnorm"]) for r in done_m])
    worst_p = worst([(i, err_of(i, ev), st["rnorm"]) for i, _, ev, st in plain])
    # the oracle's eigenvalues of the stream's small full-CI requests (tests/golden, made by
    # its script there with the oracle's CPU Davidson at residual 1e-9), where the stream is
    # the one the fixture was made for and the request's spec matches
    gold = {}
    gpath = os.path.join(ROOT, "tests", "golden", "stream%d_energies.json" % seed)
    if rate and len(stream) == 64 and os.path.exists(gpath):
        for g in json.load(open(gpath))["requests"]:
            i = g["request"]
            if stream[i][1] == g["spec"] and stream[i][2] == g["seed"]:
                gold[i] = g["energies"]

    def err_gold(items):
        errs = [max(abs(ev[k] - gold[i][k]) for k in range(len(ev))) for i, ev in items
                if i in gold]
        return {"requests": len(errs), "max_abs_err_Ha": max(errs) if errs else None}
    gold_m = err_gold([(r.idx, r.energies) for r in done_m])
    gold_b = err_gold([(r.idx, r.energies) for r in done_b])
    gold_p = err_gold([(i, ev) for i, _, ev, _ in plain])
    if gold:
        log("against the oracle's energies (%d requests): mixed %.2e, served fp64 %.2e, plain "
            "%.2e Ha" % (len(gold), gold_m["max_abs_err_Ha"], gold_b["max_abs_err_Ha"],
                         gold_p["max_abs_err_Ha"]))
    lat_m = [r.t_done - r.arrival for r in done_m]
    lat_b = [r.t_done - r.arrival for r in done_b]
    lat_p = [lat for _, lat, _, _ in plain]
    # CPU baseline work: the full-CI requests (the reference's CPU path cannot take the
    # MRCI-SD space: an O(N^2) connectivity map over 6e8 determinants)
    fci_plain = [(i, st) for i, _, _, st in plain if stream[i][1]["norbs"] <= LANE_NORBS]
    elements = sum(sum(fci_pairs(stream[i][1])) * st["nsigma64"] for i, st in fci_plain)
    return {
        "value": len(done_m) / wall_m, "ms_per_step": 1e3 * wall_m / len(done_m),
        "config": {"workload": "%s of %d requests (seed %d): full CI n~U{8..14}, "
                               "alpha=beta~U{3..5}, roots~U{1..4}%s; residual 1e-6; "
                               "200 GB admission budget, 256 slots"
                               % ("Poisson stream at %g requests/s" % rate if rate else
                                  "batch present at t=0", len(stream), seed,
                                  (", %d of them (%.0f%%) MRCI-SD 16e/%do (4 docc + "
                                   "CAS(8e,8o) + %d external, %d determinants) with "
                                   "roots~U{1..3} on the serving lane"
                                   % (nmrci, 100.0 * nmrci / len(stream),
                                      12 + args.mrci_nvirt, args.mrci_nvirt,
                                      max(n for (_, s, _), n in zip(stream, ndets)
                                          if s["norbs"] > LANE_NORBS))) if nmrci else
                                  " (no MRCI-SD share)"),
                   "mrci_requests": nmrci,
                   "requests": len(stream), "determinants_total": sum(ndets),
                   "largest_request": max(ndets), "roots_total": sum(
                       s["nroots"] for _, s, _ in stream)},
        "mixed": {"solves_per_s": len(done_m) / wall_m, "p50_s": percentile(lat_m, 0.5),
                  "p99_s": percentile(lat_m, 0.99), "max_abs_err_Ha": err_m,
                  "converged": sum(1 for r in done_m if r.stats["done"] == 1),
                  "sigma_s": sig_m, "step_phases_s": phases.get(J.PREC_MIXED),
                  "tuned_tiles_fp64_fp32": tiles.get(J.PREC_MIXED),
                  "worst": worst_m, "vs_oracle": gold_m,
                  "path": "served: continuous batching, mixed precision"},
        "plain_fp64": {"solves_per_s": len(plain) / wall_p, "p50_s": percentile(lat_p, 0.5),
                       "p99_s": percentile(lat_p, 0.99), "max_abs_err_Ha": err_p,
                       "worst": worst_p, "vs_oracle": gold_p,
                       "path": "one request at a time in arrival order, FP64"},
        "served_fp64": {"solves_per_s": len(done_b) / wall_b, "p50_s": percentile(lat_b, 0.5),
                        "p99_s": percentile(lat_b, 0.99), "max_abs_err_Ha": err_b,
                        "sigma_s": sig_b, "step_phases_s": phases.get(J.PREC_FP64),
                        "vs_oracle": gold_b,
                        "path": "served: continuous batching, FP64"},
        "reference_solves": {"restol": 1e-8, "largest_final_residual": ref_r,
                             "what": "FP64 GPU solve of every request, 2 buffer roots "
                                     "(max_abs_err_Ha); vs_oracle: the oracle's CPU Davidson "
                                     "energies of the requests of at most 30,000 "
                                     "determinants (tests/golden/stream%d_energies.json)"
                                     % seed},
        "mixed_over_plain": wall_p / wall_m,
        "_cpu": (len(fci_plain), elements),
        "_big": (stream[big][1], stream[big][2]),
        "_mrci": next(((sp_, sd_) for _, sp_, sd_ in stream if sp_["norbs"] > LANE_NORBS), None),
    }


def bench_mrci(J, ctx, args):
    """MRCI-SD 16 electrons in 60 orbitals (4 doubly occupied + CAS(8e,8o) + 48 external,
    excitation level 2, MRCI-SD rule: 603,651,972 determinants), synthetic integrals of seed
    1 (n = 60): plain FP64 solve and/or mixed precision (FP32 sigma until the residual is
    below 1e-4, then FP64) of the lowest args.roots roots, residual 1e-6 and |r|^2/gap below
    1e-9, matrix-free table sigma.  With both precisions the per-root |dE| of the mixed solve
    against the FP64 one is reported (the FP64 GPU solve is this configuration's accuracy
    reference).  The sigma kernel's rate is reported as connected entries per second (N
    times the mean connections of 64 sampled rows, per vector) and as compulsory HBM bytes
    N*(2*k*8 + 8) over the launch time."""
    nv = args.mrci_nvirt
    n = 12 + nv
    m1, m2 = J.synth_integrals(n, 1)
    t0 = time.perf_counter()
    sp = J.Space(ctx, 8, 8, n, 4, 8, 2, J.RULE_MRCISD)
    h = J.Hamiltonian(sp, m1, m2, J.DIRECT_TABLES)
    ctx.sync()
    setup_s = time.perf_counter() - t0
    N = sp.ndets
    log("MRCI-SD 16e/%do: %d determinants, tables %.1f s" % (n, N, setup_s))
    import numpy as np
    rows = np.random.default_rng(0).integers(0, N, size=64)
    cbar = float(np.mean([len(c) for c, _, _, _ in h.rows(rows)]))
    from jayci_amd.serve import lane_solve_params
    nr = args.roots
    # the subspace as large as the device's free memory allows (a margin of 6 GiB kept;
    # the diagonal and the excitation tables, 128 bytes per determinant in the lane's
    # estimate, are allocated already)
    free_b, total_b = J.device_memory(ctx.device)
    kw = lane_solve_params({"nroots": nr}, N, 1e-6, budget=free_b - (6 << 30) + 128 * N,
                           kfac=args.mrci_kfac)
    nt = kw["nroots"] + kw["nbuffer"]
    log("device memory: %.1f GB free of %.1f GB; krymax %d" % (free_b / 1e9, total_b / 1e9,
                                                              kw["krymax"]))
    kw.update(maxiter=args.mrci_maxiter, vectors=False, check=False)
    sig = {}
    for k in (sorted({1, nt}) if args.mrci_sigma_bench else []):
        for meth, name in ((J.SIGMA_FAST64, "fp64"), (J.SIGMA_FAST32, "fp32")):
            ms = h.bench_sigma(meth, k, 1)
            sig["%s_k%d" % (name, k)] = {
                "ms": ms, "entries_per_s": N * cbar * k / (ms * 1e-3),
                "compulsory_GBps": N * (2 * k * (8 if name == "fp64" else 4) + 8) / (ms * 1e-3) / 1e9}
            log("table sigma %s, %d vector(s): %.0f ms" % (name, k, ms))
    precs = {"fp64": [J.PREC_FP64], "mixed": [J.PREC_MIXED],
             "both": [J.PREC_FP64, J.PREC_MIXED]}[args.precision]
    out = {}
    for prec in precs:
        ctx.sync()
        t = time.perf_counter()
        with c_stdout_to_stderr():
            ev, _, st = h.solve(precision=prec, verbose=True, **kw)
            ctx.sync()
        wall = time.perf_counter() - t
        key = "fp64" if prec == J.PREC_FP64 else "mixed"
        nsig = st["nsigma64"] + st["nsigma32"]
        out[key] = {"seconds": wall, "energies": [float(e) for e in ev], "rc": st["rc"],
                    "sigma_vector_seconds": st["sigma_seconds"] / nsig if nsig else None,
                    "converged": st["converged"], "iterations": st["iterations"],
                    "nsigma64": st["nsigma64"], "nsigma32": st["nsigma32"],
                    "rnorm": st["rnorm"], "sigma_seconds": st["sigma_seconds"]}
        log("%s solve: %.1f s, %d iterations, energies %s, rnorm %s"
            % (key, wall, st["iterations"], ev, ["%.2e" % r for r in st["rnorm"]]))
    h.close()
    sp.close()
    res = {"config": {"workload": "MRCI-SD 16e/%do (4 docc + CAS(8e,8o) + %d external, xlvl 2, "
                                  "MRCI-SD rule), %d determinants, %d roots, residual 1e-6, "
                                  "integrals seed 1" % (n, nv, N, nr),
                      "ndets": N, "setup_s": setup_s, "connections_per_row": cbar},
           "sigma_table_kernel": sig}
    res.update(out)
    if "fp64" in out and "mixed" in out:
        res["abs_err_mixed_vs_fp64_Ha"] = [abs(a - b) for a, b in zip(out["mixed"]["energies"],
                                                                      out["fp64"]["energies"])]
        res["mixed_over_plain"] = out["fp64"]["seconds"] / out["mixed"]["seconds"]
    main_key = "mixed" if "mixed" in out else "fp64"
    res["value"] = 1.0 / out[main_key]["seconds"]
    res["ms_per_step"] = 1e3 * out[main_key]["seconds"]
    return res


def dropin_leg(J, wl, m1, m2, gold):
    """The reference's driver sequence through the drop-in entry points on the same space
    (source/execute_ci_calculation.c:206-237): citrunc_rtnlist -> generate_cimap (the
    reference's section lists, from the GPU screen) -> compute_hdgls -> dvdalg (predr 1,
    krymin 4, krymax 16, restol 1e-6; the list comes from citrunc_rtnlist, so dvdalg runs the
    handle path's sigma) -> deallocate_cimap, each phase timed."""
    import ctypes
    import numpy as np
    from jayci_amd import dropin
    L = J.lib()
    c = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    out = {}
    with c_stdout_to_stderr():
        t = time.perf_counter()
        dl, info = dropin.citrunc_rtnlist(wl["na"], wl["nb"], wl["norbs"], 0, wl["ndocc"],
                                          wl["nactv"], 0, wl["xlvl"])
        out["citrunc_rtnlist_s"] = time.perf_counter() - t
        N = info["ndets"]
        ninto = wl["ndocc"] + wl["nactv"]
        t = time.perf_counter()
        rc, hmap = dropin.generate_cimap(dl, N, ninto)
        out["generate_cimap_s"] = time.perf_counter() - t
        hd = np.zeros(N)
        t = time.perf_counter()
        L.compute_hdgls(dl, N, c(m1), c(m2), wl["na"], wl["nb"], c(hd), ninto)
        out["compute_hdgls_s"] = time.perf_counter() - t
        civ = np.zeros((1, N))
        cival = np.zeros(1)
        t = time.perf_counter()
        rc = L.dvdalg(dl, N, c(m1), c(m2), wl["na"], wl["nb"], c(hd), ninto, J.SYNTH_ENUC, 200,
                      4, 16, 1, 1e-6, hmap, dropin.row_pointers(civ), c(cival), 1, 0,
                      wl["norbs"], 64)
        out["dvdalg_s"] = time.perf_counter() - t
        t = time.perf_counter()
        dropin.deallocate_cimap(hmap, N)
        out["deallocate_cimap_s"] = time.perf_counter() - t
        dropin.free(dl)
    out["rc"] = rc
    out["energy"] = float(cival[0])
    out["abs_err_vs_golden_Ha"] = abs(float(cival[0]) - gold[0])
    out["path"] = ("reference entry points (libjayci_gpu drop-in): section lists built on the "
                   "host from the GPU screen; dvdalg on the handle path's sigma")
    log("drop-in sequence: map %.1f s, dvdalg %.3f s, |dE| %.1e"
        % (out["generate_cimap_s"], out["dvdalg_s"], out["abs_err_vs_golden_Ha"]))
    return out


def bench_fci12(J, ctx, args):
    from jayci_amd.workloads import WORKLOADS
    wl = WORKLOADS["fci12"]
    m1, m2 = J.synth_integrals(wl["norbs"], wl["seed"])
    sp = J.Space(ctx, wl["na"], wl["nb"], wl["norbs"], wl["ndocc"], wl["nactv"], wl["xlvl"],
                 wl["rule"])
    t0 = time.time()
    h = J.Hamiltonian(sp, m1, m2, J.PAIRS_FP64 | J.PAIRS_FP32 | J.DIRECT_TABLES)
    setup_s = time.time() - t0
    b64, b32, tune, width = h.autotune(nvec=1, reps=5)
    N = sp.ndets
    log("setup %.2f s, autotune %s -> %d / %d (W=%d)" % (setup_s, tune, b64, b32, width))
    gold = json.load(open(os.path.join(ROOT, "tests", "golden", "cfg0_energies.json")))["energies"]
    kw = dict(nroots=wl["nroots"], restol=1e-6, krymin=4, krymax=16, refdim=64,
              totfrze=J.SYNTH_ENUC, vectors=False)

    def run(prec, k):
        lat, err, st = [], [], None
        for _ in range(k):
            ctx.sync()
            t = time.perf_counter()
            ev, _, st = h.solve(precision=prec, **kw)
            ctx.sync()
            lat.append(time.perf_counter() - t)
            err.append(max(abs(ev[r] - gold[r]) for r in range(wl["nroots"])))
        return lat, err, st
    run(J.PREC_MIXED, args.warmup)
    run(J.PREC_FP64, args.warmup)
    lm, em, sm = run(J.PREC_MIXED, args.steps)
    lp, ep, spl = run(J.PREC_FP64, args.steps)
    h.close()
    sp.close()
    spec = dict(na=wl["na"], nb=wl["nb"], norbs=wl["norbs"], ndocc=wl["ndocc"],
                nactv=wl["nactv"], xlvl=wl["xlvl"], rule=wl["rule"])
    drop = dropin_leg(J, wl, m1, m2, gold) if args.dropin else None
    return {
        "dropin": drop,
        "value": args.steps / sum(lm), "ms_per_step": 1e3 * sum(lm) / args.steps,
        "config": {"workload": wl["desc"], "ndets": N,
                   "setup_s": round(setup_s, 3), "tile_width": width,
                   "sigma_fp64": METHOD_NAMES.get(b64), "sigma_fp32": METHOD_NAMES.get(b32),
                   "sigma_ms": {METHOD_NAMES[m]: round(v, 4) for m, v in tune.items()}},
        "mixed": {"solves_per_s": args.steps / sum(lm), "p50_s": percentile(lm, 0.5),
                  "p99_s": percentile(lm, 0.99), "max_abs_err_Ha": max(em),
                  "nsigma32": sm["nsigma32"], "nsigma64": sm["nsigma64"]},
        "plain_fp64": {"solves_per_s": args.steps / sum(lp), "p50_s": percentile(lp, 0.5),
                       "p99_s": percentile(lp, 0.99), "max_abs_err_Ha": max(ep),
                       "nsigma64": spl["nsigma64"]},
        "mixed_over_plain": sum(lp) / sum(lm),
        "_cpu": (1, sum(fci_pairs(spec)) * spl["nsigma64"]),
        "_big": (spec, wl["seed"]),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--workload", default="served",
                    choices=["served", "batched", "fci12", "mrci"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--dropin", action="store_true",
                    help="fci12: also time the drop-in entry points' driver sequence")
    ap.add_argument("--roots", type=int, default=3, help="mrci: roots to solve for")
    ap.add_argument("--precision", default="both", choices=["fp64", "mixed", "both"],
                    help="mrci: which solves to run")
    ap.add_argument("--mrci-nvirt", type=int, default=None,
                    help="external orbitals of the MRCI-SD space: mrci workload (default 48, "
                         "the configuration's), served stream's share (default 8: 16e/20o, "
                         "17,909,892 determinants, whose solves take seconds; at 48 one takes "
                         "tens of minutes)")
    ap.add_argument("--mrci-maxiter", type=int, default=200)
    ap.add_argument("--mrci-kfac", type=int, default=6,
                    help="mrci: subspace of this many times the tracked pairs, as memory "
                         "allows (twice at the least)")
    ap.add_argument("--mrci-sigma-bench", action="store_true",
                    help="mrci: also time single sigma launches of 1 and the block's vectors")
    ap.add_argument("--mrci-share", type=float, default=0.1,
                    help="served: share of MRCI-SD 16e/(12 + mrci-nvirt)o requests in the "
                         "stream")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (there is no multi-GPU path); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.mrci_nvirt is None:
        args.mrci_nvirt = 48 if args.workload == "mrci" else 8
    if args.steps is None:
        args.steps = {"served": 64, "batched": 256, "fci12": 5, "mrci": 1}[args.workload]
    if args.warmup is None:
        args.warmup = {"served": 4, "batched": 4, "fci12": 1, "mrci": 0}[args.workload]

    import jayci_amd as J
    ctx = J.Context(0)
    if args.workload == "mrci":
        res = bench_mrci(J, ctx, args)
        line = {"metric": METRIC, "value": res.pop("value"), "unit": "solves/s", "n_gpus": 1,
                "steps": 1, "warmup": 0, "ms_per_step": res.pop("ms_per_step"),
                "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                "dtype": "f64" if args.precision == "fp64" else "f32 sigma -> f64 refinement",
                "data": "synthetic integrals (SplitMix64/Box-Muller), seed 1"}
        line.update(res)
        print(json.dumps(line))
        return
    if args.workload == "fci12":
        res = bench_fci12(J, ctx, args)
    elif args.workload == "served":
        res = bench_stream(J, ctx, args, 200.0, 400)
    else:
        res = bench_stream(J, ctx, args, 0.0, 300)
    work = res.pop("_cpu")
    bspec, bseed = res.pop("_big")
    from jayci_amd.workloads import WORKLOADS
    wl = WORKLOADS["fci12"]
    plain_roof = roofline_of(J, ctx, {k: wl[k] for k in ("na", "nb", "norbs", "ndocc", "nactv",
                                                        "xlvl", "rule")}, wl["seed"])
    log("pair-list sigma %.1f GB/s (%.3f of HBM peak)" % (plain_roof["achieved"], plain_roof["frac"]))
    if plain_roof.get("dense_string_fp64_ms"):
        log("same sigma by the dense-string path: %.3f ms against %.3f ms for the pair bytes at "
            "HBM peak" % (plain_roof["dense_string_fp64_ms"], plain_roof["roofline_ms_at_peak"]))
    dense = roofline_dense(J, ctx, bspec, bseed)
    roof, roof32 = dense[64], dense[32]
    log("dense-string sigma FP64 %.1f TFLOP/s (%.3f), FP32 %.1f TFLOP/s (%.3f)"
        % (dense[64]["achieved"], dense[64]["frac"], dense[32]["achieved"], dense[32]["frac"]))
    mrci = res.pop("_mrci", None)
    table = None
    if mrci is not None:
        # with the MRCI-SD share the stream's time is the lane's table sigma: it is the
        # dominant kernel, the dense-string sigma of the full-CI requests is reported beside it
        table = roofline_table(J, ctx, mrci[0], mrci[1])
        roof, roof32 = table[64], table[32]
        log("table sigma FP64 %.3f ms per pass, %.2f GB/s compulsory (%.5f of HBM peak), "
            "%.3g entries/s" % (table[64]["ms_per_launch"], table[64]["achieved"],
                                table[64]["frac"], table[64]["connected_entries_per_s"]))
    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(J, work)
        except Exception as e:  # reported, not fatal
            cpu = {"value": None, "unit": "solves/s", "cores": None, "kind": "port",
                   "sample": "failed: %s" % e}
    line = {"metric": METRIC, "value": res["value"], "unit": "solves/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": res["ms_per_step"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32 sigma -> f64 refinement (value); f64 (plain_fp64)",
            "data": "synthetic integrals (SplitMix64/Box-Muller), seeded",
            "config": res["config"], "mixed": res["mixed"], "plain_fp64": res["plain_fp64"]}
    if "served_fp64" in res:
        line["served_fp64"] = res["served_fp64"]
    if res.get("dropin"):
        line["dropin"] = res["dropin"]
    line.update({"mixed_over_plain": res["mixed_over_plain"], "roofline": roof,
                 "roofline_fp32": roof32, "roofline_pair_list": plain_roof,
                 "cpu_baseline": cpu})
    if table is not None:
        line.update({"roofline_dense": dense[64], "roofline_dense_fp32": dense[32]})
    print(json.dumps(line))


if __name__ == "__main__":
    main()
