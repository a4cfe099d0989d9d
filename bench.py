"""Benchmark: Davidson solves/s and latency, mixed precision vs plain FP64, one GPU.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload fci12]

A step is one complete Davidson solve (lowest root(s) to residual 1e-6 and 1e-8 Ha of the
golden energy) on synthetic integrals resident on the GPU.  K solves of the mixed
(FP32 sigma -> FP64 refinement) path are timed, then K solves of the plain FP64 path on
the same inputs for the comparison; value = mixed solves/s.

The line also carries the roofline of the dominant kernel (the pair-list sigma, timed
with HIP events on its own stream inside the library) and the CPU baseline (the
reference's own compute_hv, compiled from its sources into oracle/_ref, timed on this
host for one sigma pass and scaled by the sigma passes of a plain solve).
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM = 8.0e12        # MI355X HBM3E, bytes/s

from jayci_amd.workloads import WORKLOADS  # noqa: E402  (plain dict module)

METHOD_NAMES = {1: "pairs64", 2: "pairs32", 3: "pairs_sym64", 4: "direct64", 5: "direct32",
                6: "fast64", 7: "fast32"}


def percentile(xs, q):
    xs = sorted(xs)
    if not xs:
        return None
    k = (len(xs) - 1) * q
    f = int(k)
    c = min(f + 1, len(xs) - 1)
    return xs[f] + (xs[c] - xs[f]) * (k - f)


def golden_energy(name):
    path = os.path.join(ROOT, "tests", "golden", "cfg0_energies.json")
    if name == "fci12" and os.path.exists(path):
        return json.load(open(path))["energies"]
    return None


def cpu_baseline(wl, m1, m2, nsigma_per_solve):
    """The reference's serial compute_hv (source/action_util.c:121, OpenMP) for one sigma
    pass over the whole space, scaled to solves/s by the sigma passes of a plain GPU
    solve.  Its connectivity map comes from this library's generate_cimap (same section
    lists as source/cimapping.c:92; checked bit-exact in the tests)."""
    import ctypes
    import numpy as np
    from oracle import ref
    if not ref.available():
        return None
    os.environ.setdefault("OMP_SCHEDULE", "dynamic")
    import jayci_amd as J
    rs = ref.RefSpace(wl["na"], wl["nb"], wl["norbs"], wl["ndocc"], wl["nactv"], wl["xlvl"])
    hmap = ref._libc.malloc(ctypes.sizeof(ref.RowMap) * rs.ndets)
    gen = J.lib().generate_cimap
    gen.restype = ctypes.c_int
    gen.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    rc = gen(ctypes.cast(rs.dlist, ctypes.c_void_p), rs.ndets, rs.ninto, hmap)
    if rc != 0:
        return None
    v = np.random.default_rng(0).standard_normal(rs.ndets)
    t0 = time.time()
    rs.compute_hv(m1, m2, v, hmap)
    t = time.time() - t0
    dealloc = J.lib().deallocate_cimap
    dealloc.argtypes = [ctypes.c_void_p, ctypes.c_int]
    dealloc(hmap, rs.ndets)
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": 1.0 / (t * nsigma_per_solve), "unit": "solves/s", "cores": cores,
            "kind": "reference",
            "sample": "one full compute_hv pass of the reference (oracle/_ref, gcc -O2 "
                      "-fopenmp) over the %d-determinant space: %.2f s; scaled by the %d "
                      "sigma applications of a plain FP64 GPU solve" % (rs.ndets, t,
                                                                        nsigma_per_solve)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="fci12", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (no multi-GPU path); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import jayci_amd as J

    wl = WORKLOADS[args.workload]
    ctx = J.Context(0)
    m1, m2 = J.synth_integrals(wl["norbs"], wl["seed"])
    sp = J.Space(ctx, wl["na"], wl["nb"], wl["norbs"], wl["ndocc"], wl["nactv"], wl["xlvl"],
                 wl["rule"])
    t0 = time.time()
    h = J.Hamiltonian(sp, m1, m2, J.PAIRS_FP64 | J.PAIRS_FP32 | J.DIRECT_TABLES)
    setup_s = time.time() - t0
    # autotuner: fastest sigma variant / tile width for this space becomes the default
    best64, best32, tune_ms, width = h.autotune(nvec=1, reps=5)
    N = sp.ndets
    gold = golden_energy(args.workload)
    solve_kw = dict(nroots=wl["nroots"], restol=1e-6, krymin=4, krymax=16, refdim=64,
                    totfrze=J.SYNTH_ENUC, vectors=False)

    def log(msg):
        print("bench: " + msg, file=sys.stderr, flush=True)

    log("setup %.2f s, N=%d, pairs=%d, autotune %s -> fp64 %d fp32 %d (W=%d)"
        % (setup_s, N, h.npairs, tune_ms, best64, best32, width))

    def run(precision, k):
        lat, errs, stats = [], [], []
        for _ in range(k):
            ctx.sync()
            t = time.perf_counter()
            ev, _, st = h.solve(precision=precision, **solve_kw)
            ctx.sync()
            lat.append(time.perf_counter() - t)
            stats.append(st)
            if gold:
                errs.append(max(abs(ev[r] - gold[r]) for r in range(wl["nroots"])))
            log("solve prec=%d %.4f s E0 %.12f iters %d" % (precision, lat[-1], ev[0],
                                                             st["iterations"]))
        return lat, errs, stats

    run(J.PREC_MIXED, args.warmup)
    run(J.PREC_FP64, args.warmup)
    lat_m, err_m, st_m = run(J.PREC_MIXED, args.steps)
    lat_p, err_p, st_p = run(J.PREC_FP64, args.steps)
    tot_m, tot_p = sum(lat_m), sum(lat_p)

    # roofline of the dominant kernel: FP64 pair-list sigma, one vector
    P_full = h.npairs
    P_u = P_full // 2
    B_sigma = P_u * (8 + 4) + N * 1 * 2 * 8 + 8 * N
    ms = h.bench_sigma(J.SIGMA_PAIRS64, 1, 20)
    achieved = B_sigma / (ms * 1e-3)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "sigma_pairs64_fci12_traffic.json")
    if os.path.exists(tpath):
        traffic = json.load(open(tpath)).get("bytes_per_launch")

    log("sigma pairs64 %.4f ms/launch" % ms)
    cpu = None
    if not args.no_cpu:
        nsig = st_p[0]["nsigma64"] if st_p else 0
        try:
            cpu = cpu_baseline(wl, m1, m2, max(nsig, 1))
        except Exception as e:  # reported, not fatal
            cpu = {"value": None, "unit": "solves/s", "cores": None, "kind": "reference",
                   "sample": "failed: %s" % e}

    line = {
        "metric": "Davidson solves/s + p99 latency at 1e-8 Ha, mixed vs plain FP64, 1 GPU",
        "value": args.steps / tot_m,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * tot_m / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 sigma -> f64 refinement (mixed); f64 (plain)",
        "data": "synthetic integrals (SplitMix64/Box-Muller, seed %d)" % wl["seed"],
        "config": {"workload": wl["desc"], "ndets": N, "pairs_full": P_full,
                   "setup_s": round(setup_s, 3),
                   "sigma_fp64": METHOD_NAMES.get(best64), "sigma_fp32": METHOD_NAMES.get(best32),
                   "tile_width": width,
                   "sigma_ms": {METHOD_NAMES[m]: round(v, 4) for m, v in tune_ms.items()}},
        "mixed": {"solves_per_s": args.steps / tot_m, "p50_s": percentile(lat_m, 0.5),
                  "p99_s": percentile(lat_m, 0.99),
                  "max_abs_err_Ha": max(err_m) if err_m else None,
                  "nsigma32": st_m[0]["nsigma32"], "nsigma64": st_m[0]["nsigma64"],
                  "iterations": st_m[0]["iterations"]},
        "plain_fp64": {"solves_per_s": args.steps / tot_p, "p50_s": percentile(lat_p, 0.5),
                       "p99_s": percentile(lat_p, 0.99),
                       "max_abs_err_Ha": max(err_p) if err_p else None,
                       "nsigma64": st_p[0]["nsigma64"], "iterations": st_p[0]["iterations"]},
        "mixed_over_plain": tot_p / tot_m,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM, "traffic": traffic,
                     "kernel": "k_spmv<double> (FP64 pair list, 1 vector)",
                     "ms_per_launch": ms, "algorithmic_bytes": B_sigma},
        "cpu_baseline": cpu,
    }
    line["roofline"]["achieved"] = achieved / 1e9
    line["roofline"]["peak"] = PEAK_HBM / 1e9
    print(json.dumps(line))


if __name__ == "__main__":
    main()
