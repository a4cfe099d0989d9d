"""Launch the sigma kernels of one workload a few times for a profiler run.

    rocprofv3 --kernel-trace --stats -d OUT -- python jayci_amd/prof_sigma.py --workload fci12
    rocprofv3 --pmc FETCH_SIZE      -d OUT -- python jayci_amd/prof_sigma.py --workload fci12

Order of launches (each preceded by one warm-up launch): a 1 GiB device copy (k_copy,
the byte-counter calibration: 1 GiB read + 1 GiB written per launch), the FP64 pair-list
sigma (k_spmv<double>), the FP32 pair-list sigma (k_spmv<float>), the matrix-free
table sigma (k_sigma_fast, autotuned over tile widths), the symmetric upper-triangle sigma
(k_spmv_sym) and the one-thread-per-row matrix-free sigma (k_sigma_generic<double>).
"""
import argparse
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import jayci_amd as J  # noqa: E402
from jayci_amd.workloads import WORKLOADS  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="fci12", choices=sorted(WORKLOADS))
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--skip-direct", action="store_true")
    a = ap.parse_args()
    wl = WORKLOADS[a.workload]
    ctx = J.Context(0)
    m1, m2 = J.synth_integrals(wl["norbs"], wl["seed"])
    sp = J.Space(ctx, wl["na"], wl["nb"], wl["norbs"], wl["ndocc"], wl["nactv"], wl["xlvl"], wl["rule"])
    print("copy 1GiB  ms %.4f" % ctx.bench_copy(1 << 30, a.reps), flush=True)
    h = J.Hamiltonian(sp, m1, m2, J.PAIRS_FP64 | J.PAIRS_FP32 | J.DIRECT_TABLES)
    print("N %d pairs %d" % (sp.ndets, h.npairs), flush=True)
    print("spmv64     ms %.4f" % h.bench_sigma(J.SIGMA_PAIRS64, 1, a.reps), flush=True)
    print("spmv32     ms %.4f" % h.bench_sigma(J.SIGMA_PAIRS32, 1, a.reps), flush=True)
    b64, b32, t, w = h.autotune(1, a.reps)
    print("autotune   %s best64 %d best32 %d W %d" % (t, b64, b32, w), flush=True)
    print("fast64     ms %.4f" % h.bench_sigma(J.SIGMA_FAST64, 1, a.reps), flush=True)
    print("fast32     ms %.4f" % h.bench_sigma(J.SIGMA_FAST32, 1, a.reps), flush=True)
    print("fast64 k=3 ms %.4f" % h.bench_sigma(J.SIGMA_FAST64, 3, a.reps), flush=True)
    print("spmv64 k=3 ms %.4f" % h.bench_sigma(J.SIGMA_PAIRS64, 3, a.reps), flush=True)
    h.close()
    hu = J.Hamiltonian(sp, m1, m2, J.PAIRS_FP64 | J.PAIRS_UPPER)
    print("spmv_sym   ms %.4f" % hu.bench_sigma(J.SIGMA_PAIRS_SYM64, 1, a.reps), flush=True)
    if not a.skip_direct:
        print("direct64   ms %.4f" % hu.bench_sigma(J.SIGMA_DIRECT64, 1, 2), flush=True)
    hu.close()


if __name__ == "__main__":
    main()
