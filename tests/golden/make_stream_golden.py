"""Oracle eigenvalues of the small full-CI requests of the default served stream (64
requests at 200/s, seed 400: `jayci_amd.serve.make_stream(64, 200.0, 400)`; the requests
of at most 30,000 determinants, 31 of them), each on its own synthetic integrals, computed
with the oracle's CPU Davidson (FP64, residual 1e-9, the requested roots).  The MRCI-SD
share of the stream replaces some requests by others; the fixture is keyed by request index
and spec, so a comparison takes only the requests whose spec matches.
Run in the build container:  python tests/golden/make_stream_golden.py
Writes tests/golden/stream400_energies.json."""
import json
import os
import sys
import time
from math import comb

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import jayci_amd as J  # noqa: E402  (integral generator and stream, host code only)
from jayci_amd import serve  # noqa: E402
import oracle  # noqa: E402

MAXN = 30000
out = {"stream": "serve.make_stream(64, 200.0, 400)", "enuc": J.SYNTH_ENUC, "restol": 1e-9,
       "method": "oracle/jayci_oracle.c orc_davidson (source/davidson.c control flow)",
       "requests": []}
t00 = time.time()
for idx, (_, s, seed) in enumerate(serve.make_stream(64, 200.0, 400)):
    N = comb(s["norbs"], s["na"]) * comb(s["norbs"], s["nb"])
    if N > MAXN:
        continue
    t0 = time.time()
    m1, m2 = J.synth_integrals(s["norbs"], seed)
    sp = oracle.Space(s["na"], s["nb"], s["norbs"], 0, s["norbs"], 0, 1)
    nr = s["nroots"]
    rc, ev, _, rn, ns = sp.davidson(m1, m2, totfrze=J.SYNTH_ENUC, maxiter=600,
                                    krymin=min(nr + 3, N), krymax=min(max(8 * (nr + 2), 24), N),
                                    nroots=nr, restol=1e-9, refdim=64, nthreads=8)
    out["requests"].append({"request": idx, "spec": s, "seed": seed, "ndets": N,
                            "energies": list(map(float, ev)), "rc": rc, "rnorm_last": rn,
                            "nsigma": ns, "seconds": time.time() - t0})
    print(idx, N, nr, rc, ["%.12f" % e for e in ev], "%.1f s" % (time.time() - t0), flush=True)
out["seconds"] = time.time() - t00
json.dump(out, open(os.path.join(ROOT, "tests", "golden", "stream400_energies.json"), "w"),
          indent=1)
