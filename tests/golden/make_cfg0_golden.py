"""Golden eigenvalues of the FCI 4a4b / 12-orbital space (245,025 determinants) on the
synthetic integrals of seed 100, computed with the oracle's CPU Davidson (FP64, residual
1e-9).  Run in the build container:  python tests/golden/make_cfg0_golden.py
Writes tests/golden/cfg0_energies.json."""
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import jayci_amd as J  # noqa: E402  (integral generator only; host code)
import oracle  # noqa: E402

SEED = 100
t0 = time.time()
m1, m2 = J.synth_integrals(12, SEED)
sp = oracle.Space(4, 4, 12, 0, 12, 0, 1)
rc, ev, vec, rn, ns = sp.davidson(m1, m2, totfrze=J.SYNTH_ENUC, maxiter=400, krymin=4, krymax=16,
                                  nroots=3, restol=1e-9, refdim=64, nthreads=8)
out = {"space": "FCI 4a4b/12 (245025 dets)", "seed": SEED, "enuc": J.SYNTH_ENUC,
       "energies": list(map(float, ev)), "rc": rc, "rnorm_last": rn, "nsigma": ns,
       "restol": 1e-9, "seconds": time.time() - t0,
       "method": "oracle/jayci_oracle.c orc_davidson (source/davidson.c control flow)"}
json.dump(out, open(os.path.join(ROOT, "tests", "golden", "cfg0_energies.json"), "w"), indent=1)
print(out)
