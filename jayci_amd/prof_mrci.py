"""One matrix-free sigma pass on the MRCI-SD space of 16 electrons in 60 orbitals (4 doubly
occupied + CAS(8e,8o) + 48 external, excitation level 2; 603,651,972 determinants), for a
profiler run:

    rocprofv3 --kernel-trace --stats -d OUT -- python jayci_amd/prof_mrci.py [--fp32]

Builds the string tables, the diagonal and the per-string excitation tables on the device,
then applies the table kernel (k_sigma_fast) once to a seeded random vector through the
host-array entry point and prints the wall times of each stage; the kernel's own duration
is read from the profiler.  Also prints the connections per row over a sample of rows (the
pairs/s figure) and the compulsory HBM bytes of a pass, N*(8 + 8 + 8): vector read, sigma
written, diagonal read.
"""
import argparse
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import numpy as np  # noqa: E402

import jayci_amd as J  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--fp32", action="store_true")
    ap.add_argument("--seed", type=int, default=1)
    a = ap.parse_args()
    ctx = J.Context(0)
    t0 = time.time()
    m1, m2 = J.synth_integrals(60, a.seed)
    sp = J.Space(ctx, 8, 8, 60, 4, 8, 2, 1)
    N = sp.ndets
    print("space: %d determinants, %d blocks (%.1f s)" % (N, len(sp.blocks()), time.time() - t0),
          flush=True)
    t0 = time.time()
    h = J.Hamiltonian(sp, m1, m2, J.DIRECT_TABLES)
    ctx.sync()
    print("diagonal + excitation tables: %.1f s" % (time.time() - t0), flush=True)
    rows = np.random.default_rng(0).integers(0, N, size=64)
    cbar = float(np.mean([len(c) for c, _, _, _ in h.rows(rows)]))
    print("connections per row (64 sampled rows): %.0f -> %.3g connected entries per pass"
          % (cbar, N * cbar), flush=True)
    t0 = time.time()
    x = np.random.default_rng(a.seed).standard_normal(N)
    print("random vector: %.1f s" % (time.time() - t0), flush=True)
    meth = J.SIGMA_FAST32 if a.fp32 else J.SIGMA_FAST64
    t0 = time.time()
    y = h.sigma(x, method=meth)
    t = time.time() - t0
    print("sigma %s incl. host transfers: %.1f s; |sigma| = %.6e; compulsory bytes %.3g"
          % ("FP32" if a.fp32 else "FP64", t, float(np.linalg.norm(y)), N * 24.0), flush=True)


if __name__ == "__main__":
    main()
