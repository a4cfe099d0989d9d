"""Launch the dense-string full-CI sigma a few times for a profiler run.

    rocprofv3 --kernel-trace --stats -d OUT -- python jayci_amd/prof_dense.py
    rocprofv3 --pmc FETCH_SIZE      -d OUT -- python jayci_amd/prof_dense.py

Space: full CI 5 alpha + 5 beta electrons in 14 orbitals (4,008,004 determinants, the
largest request the served stream draws), synthetic integrals seed 7.  Launches, each
preceded by one warm-up: a 1 GiB device copy (k_copy, byte-counter calibration), then
`--reps` FP64 sigmas (k_gE<double,32>, k_gss<double>, k_gga<double>) and `--reps` FP32 sigmas
(k_gcv, k_gE<float,32>, k_gss<float>, k_gga<float>) of one vector, then the same with
--nvec vectors.
"""
import argparse
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import jayci_amd as J  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--norbs", type=int, default=14)
    ap.add_argument("--nelec", type=int, default=5)
    ap.add_argument("--nvec", type=int, default=4)
    a = ap.parse_args()
    ctx = J.Context(0)
    m1, m2 = J.synth_integrals(a.norbs, 7)
    sp = J.Space(ctx, a.nelec, a.nelec, a.norbs)
    print("copy 1GiB  ms %.4f" % ctx.bench_copy(1 << 30, a.reps), flush=True)
    h = J.Hamiltonian(sp, m1, m2, J.DIRECT_TABLES)
    print("N %d" % sp.ndets, flush=True)
    for k in (1, a.nvec):
        print("gemm64 k=%d ms %.4f" % (k, h.bench_sigma(J.SIGMA_GEMM64, k, a.reps)), flush=True)
        print("gemm32 k=%d ms %.4f" % (k, h.bench_sigma(J.SIGMA_GEMM32, k, a.reps)), flush=True)
    h.close()
    sp.close()


if __name__ == "__main__":
    main()
