"""Times the reference's serial CPU path (oracle/_ref, the reference sources compiled where
they lie by oracle/Makefile) on full CI 4a4b / 12 orbitals (245,025 determinants, seed 100
integrals), and the oracle's own sigma on the same instance, in this container:

    python oracle/time_reference.py [--threads 8]

Phases, as source/execute_ci_calculation.c:206-237 runs them: citrunc_rtnlist
(source/citruncate.c:123), generate_cimap (source/cimapping.c:92), compute_hdgls
(source/action_util.c:106) and one compute_hv pass (source/action_util.c:121).  The
reference's dvdalg is not built (it needs LAPACK, which this image does not provide), so
the Davidson total is the sigma pass time times the sigma count of the oracle's
restatement of dvdalg (tests/golden/cfg0_energies.json, residual 1e-9), stated as such.

The ratio reference-sigma / oracle-sigma calibrates the oracle, which is what is timed on
machines that must not receive reference code.  Writes profiles/cpu_reference_fci12.json.
TEST / MEASUREMENT INFRASTRUCTURE ONLY.
"""
import argparse
import json
import os
import platform
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    a = ap.parse_args()
    os.environ["OMP_NUM_THREADS"] = str(a.threads)
    os.environ.setdefault("OMP_SCHEDULE", "dynamic")
    import oracle
    from oracle import ref
    import jayci_amd as J          # integral generator (host code only)
    if not ref.available():
        sys.exit("oracle/_ref is not built (needs the reference tree)")
    m1, m2 = J.synth_integrals(12, 100)
    out = {"instance": "FCI 4a4b/12, 245,025 determinants, synthetic integrals seed 100",
           "threads": a.threads, "cpu": cpu_model(), "omp_schedule": os.environ["OMP_SCHEDULE"],
           "reference_build": "gcc -O2 -fopenmp -std=gnu11 -fcommon (oracle/Makefile, target ref)",
           "oracle_build": "gcc -O2 -fopenmp (oracle/Makefile, liboracle.so)"}
    t = time.time()
    rs = ref.RefSpace(4, 4, 12, 0, 12, 0)
    out["citrunc_rtnlist_s"] = time.time() - t
    N = rs.ndets
    print("ndets", N, flush=True)
    t = time.time()
    hmap, _ = None, None
    L = ref.lib()
    import ctypes
    hmap = ref._libc.malloc(ctypes.sizeof(ref.RowMap) * N)
    if L.generate_cimap(rs.dlist, N, rs.ninto, hmap) != 0:
        sys.exit("generate_cimap failed")
    out["generate_cimap_s"] = time.time() - t
    print("generate_cimap %.1f s" % out["generate_cimap_s"], flush=True)
    t = time.time()
    rs.hdgls(m1, m2)
    out["compute_hdgls_s"] = time.time() - t
    v = np.random.default_rng(0).standard_normal(N)
    t = time.time()
    c_ref = rs.compute_hv(m1, m2, v, hmap)
    out["compute_hv_pass_s"] = time.time() - t
    print("compute_hv %.2f s" % out["compute_hv_pass_s"], flush=True)
    L.deallocate_cimap(hmap, N)
    osp = oracle.Space(4, 4, 12)
    osp.sigma(m1, m2, v, nthreads=a.threads)           # warm
    ts = []
    for _ in range(3):
        t = time.time()
        c_orc = osp.sigma(m1, m2, v, nthreads=a.threads)
        ts.append(time.time() - t)
    out["oracle_sigma_pass_s"] = min(ts)
    out["sigma_ratio_reference_over_oracle"] = out["compute_hv_pass_s"] / out["oracle_sigma_pass_s"]
    # the two sigmas differ only by the serial reference's same-spin double rule
    out["sigma_rel_diff"] = float(np.abs(c_ref - c_orc).max() / np.abs(c_orc).max())
    gold = json.load(open(os.path.join(ROOT, "tests", "golden", "cfg0_energies.json")))
    out["oracle_dvdalg_sigma_count_3roots_1e-9"] = gold["nsigma"]
    out["davidson_estimate_s"] = gold["nsigma"] * out["compute_hv_pass_s"]
    out["end_to_end_estimate_s"] = (out["citrunc_rtnlist_s"] + out["generate_cimap_s"] +
                                    out["compute_hdgls_s"] + out["davidson_estimate_s"])
    out["pairs_unique"] = 174457800
    out["elements_per_pass"] = 174457800 + N
    out["reference_elements_per_s"] = out["elements_per_pass"] / out["compute_hv_pass_s"]
    out["oracle_elements_per_s"] = out["elements_per_pass"] / out["oracle_sigma_pass_s"]
    path = os.path.join(ROOT, "profiles", "cpu_reference_fci12.json")
    json.dump(out, open(path, "w"), indent=1)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
