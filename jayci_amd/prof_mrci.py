"""One matrix-free sigma pass on the MRCI-SD space of 16 electrons in 60 orbitals (4 doubly
occupied + CAS(8e,8o) + 48 external, excitation level 2; 603,651,972 determinants), for a
profiler run:

    rocprofv3 --kernel-trace --stats -d OUT -- python jayci_amd/prof_mrci.py [--fp32] [--nvec K]
    rocprofv3 --pmc FETCH_SIZE -d OUT -- python jayci_amd/prof_mrci.py

Runs a 1 GiB device copy (the byte counters' calibration), then builds the string tables,
the diagonal and the per-string excitation tables on the device,
then applies the table sigma (two passes, per-block tile widths: the defaults) once to K
seeded random vectors resident on the device (one untimed launch first) and prints the
wall times of each stage; the kernels' own durations are read from the profiler.  Also
prints the connections per row over a sample of rows (the entries/s figure) and the
compulsory HBM bytes of a pass, N*(2*K*b + 8) with b the bytes of a vector element: vector
read, sigma written, diagonal read.
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
    ap.add_argument("--nvec", type=int, default=1)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--nvirt", type=int, default=48)
    ap.add_argument("--xg", type=int, default=-1, help="tiles per XCD run (-1: default)")
    ap.add_argument("--cm", type=int, default=-1, help="chunk-major tile order 0/1 (-1: default)")
    ap.add_argument("--split", type=int, default=-1, help="two-pass mode 0..3 (-1: default)")
    ap.add_argument("--chunk", type=int, default=-1,
                    help="alpha singles staged together, 1/2/4 (0: per width class)")
    a = ap.parse_args()
    ctx = J.Context(0)
    # a 1 GiB device copy first (k_copy: 1 GiB read, 1 GiB written), the byte counters'
    # calibration in counter runs
    print("copy 1GiB ms %.4f" % ctx.bench_copy(1 << 30, 1), flush=True)
    t0 = time.time()
    n = 12 + a.nvirt
    m1, m2 = J.synth_integrals(n, a.seed)
    sp = J.Space(ctx, 8, 8, n, 4, 8, 2, 1)
    N = sp.ndets
    print("space: %d determinants, %d blocks (%.1f s)" % (N, len(sp.blocks()), time.time() - t0),
          flush=True)
    t0 = time.time()
    h = J.Hamiltonian(sp, m1, m2, J.DIRECT_TABLES)
    ctx.sync()
    xg = h.fast_xcd(a.xg, a.cm)
    h.fast_split(a.split)
    h.fast_tiles(-1, a.chunk)
    print("diagonal + excitation tables: %.1f s; tiles %s, split %d, XCD run, chunk-major %s"
          % (time.time() - t0, h.fast_tiles(), h.fast_split(), xg), flush=True)
    rows = np.random.default_rng(0).integers(0, N, size=64)
    cbar = float(np.mean([len(c) for c, _, _, _ in h.rows(rows)]))
    print("connections per row (64 sampled rows): %.0f -> %.3g connected entries per vector"
          % (cbar, N * cbar), flush=True)
    meth = J.SIGMA_FAST32 if a.fp32 else J.SIGMA_FAST64
    t0 = time.time()
    ms = h.bench_sigma(meth, a.nvec, 1)
    b = 4 if a.fp32 else 8
    print("sigma %s, %d vector(s): %.0f ms per launch (%.1f s with setup); %.3g entries/s; "
          "compulsory bytes %.3g (%.1f GB/s)"
          % ("FP32" if a.fp32 else "FP64", a.nvec, ms, time.time() - t0,
             N * cbar * a.nvec / (ms * 1e-3), N * (2 * a.nvec * b + 8),
             N * (2 * a.nvec * b + 8) / (ms * 1e-3) / 1e9), flush=True)


if __name__ == "__main__":
    main()
