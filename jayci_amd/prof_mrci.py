"""One matrix-free sigma pass on the MRCI-SD space of 16 electrons in 60 orbitals (4 doubly
occupied + CAS(8e,8o) + 48 external, excitation level 2; 603,651,972 determinants).

    python jayci_amd/prof_mrci.py [--fp32]

Builds the string tables, the diagonal and the per-string excitation tables on the device,
then times one sigma pass (k_sigma_fast, HIP events on the library stream) on one random
vector and reports connected pairs per second and the compulsory HBM bytes of the pass
(N*(2*8 + 8): vector read, sigma written, diagonal read).
"""
import argparse
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

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
    print("space: %d determinants, %d blocks (%.1f s)" % (sp.ndets, len(sp.blocks()), time.time() - t0),
          flush=True)
    t0 = time.time()
    h = J.Hamiltonian(sp, m1, m2, J.DIRECT_TABLES)
    ctx.sync()
    print("diagonal + excitation tables: %.1f s" % (time.time() - t0), flush=True)
    meth = J.SIGMA_FAST32 if a.fp32 else J.SIGMA_FAST64
    ms = h.bench_sigma(meth, 1, 1)
    N = sp.ndets
    # connected off-diagonal entries per row, summed over a sample of rows
    import numpy as np
    rows = np.random.default_rng(0).integers(0, N, size=64)
    counts = [len(c) for c, _, _, _ in h.rows(rows)]
    cbar = float(np.mean(counts))
    pairs = N * cbar
    comp = N * (2 * 8 + 8)
    print("sigma %s: %.1f ms; sampled connections per row %.0f -> %.3g connected entries, "
          "%.3g entries/s; compulsory bytes %.3g (%.1f GB/s)"
          % ("FP32" if a.fp32 else "FP64", ms, cbar, pairs, pairs / (ms * 1e-3), comp,
             comp / (ms * 1e-3) / 1e9), flush=True)


if __name__ == "__main__":
    main()
