"""Write the SIFS integral-file fixtures under tests/golden/sifs/.

Each case is a text description turned into a SIFS file by oracle/_ref/sifs_ref (built by
oracle/sifs/build.sh from colib's header writer, label/dword packers and record writer
under the reference tree), plus the arrays the reference's readmoints
(source/readmoints.f90) returns for it, written as <case>.npz:

  H1       the sum of every itypea-0 array except the overlap (itypeb 0), symmetry-blocked
           lower triangle packed by rows (colib sifrsh / sifr1n, source/colib/colib8.f:2724,
           :1858);
  en       (energy(1), the sum of the frozen-core values of the last record of each
           H1 array);
  moints2  every 2-e value added at index2e(i, j, k, l) (readmoints.f90:220-248);
  dipoles  (type1n 3) the itypea-1 arrays x/y/z from the records with last = 1
           (readmoints.f90:166-205).

Run from the repository root after oracle/sifs/build.sh:
    python tests/golden/make_sifs_fixtures.py
"""
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
OUT = os.path.join(ROOT, "tests", "golden", "sifs")
HARNESS = os.path.join(ROOT, "oracle", "_ref", "sifs_ref")
sys.path.insert(0, ROOT)


def ind2(i, j):
    return (i - 1) * i // 2 + j if i >= j else (j - 1) * j // 2 + i


def write_case(name, nbft, nbpsy, ifmt, l1rec, n1max, l2rec, n2max, energy, ietype, arrays,
               two, two_type=(3, 0)):
    """arrays: list of (itypea, itypeb, fcore, [(i, j, v), ...]); two: [(i, j, k, l, v)]."""
    os.makedirs(OUT, exist_ok=True)
    spec = os.path.join(OUT, name + ".spec")
    with open(spec, "w") as f:
        f.write("%d %d %d %d %d %d %d\n" % (nbft, len(nbpsy), ifmt, l1rec, n1max, l2rec, n2max))
        f.write(" ".join(str(n) for n in nbpsy) + "\n%d\n" % len(energy))
        f.write(" ".join(repr(float(e)) for e in energy) + "\n")
        f.write(" ".join(str(t) for t in ietype) + "\n%d\n" % len(arrays))
        for ta, tb, fc, vals in arrays:
            f.write("%d %d %r %d\n" % (ta, tb, float(fc), len(vals)))
            for i, j, v in vals:
                f.write("%d %d %r\n" % (i, j, float(v)))
        f.write("%d %d %d\n" % (len(two), two_type[0], two_type[1]))
        for i, j, k, l, v in two:
            f.write("%d %d %d %d %r\n" % (i, j, k, l, float(v)))
    sifs = os.path.join(OUT, name + ".sifs")
    subprocess.run([HARNESS, spec, sifs], check=True)
    os.remove(spec)

    # what readmoints returns
    m1 = nbft * (nbft + 1) // 2
    m2 = m1 * (m1 + 1) // 2
    # symmetry-blocked packing of the H1 arrays
    off, nskp, nnskp, symb = 0, [], [], {}
    nn = 0
    for s, n in enumerate(nbpsy):
        nskp.append(off)
        nnskp.append(nn)
        for b in range(off + 1, off + n + 1):
            symb[b] = s
        off += n
        nn += n * (n + 1) // 2
    h1 = np.zeros(m1)
    fcore = 0.0
    dip = np.zeros(3 * m1)
    for ta, tb, fc, vals in arrays:
        if ta == 0 and 1 <= tb <= 6:
            for i, j, v in vals:
                a, b = max(i, j), min(i, j)
                s = symb[a]
                assert symb[b] == s
                a -= nskp[s]
                b -= nskp[s]
                h1[nnskp[s] + a * (a - 1) // 2 + b - 1] += v
            fcore += fc
    g = np.zeros(m2)
    for i, j, k, l, v in two:
        g[ind2(ind2(i, j), ind2(k, l)) - 1] += v
    np.savez(os.path.join(OUT, name + ".npz"), h1=h1, moints2=g,
             en=np.array([energy[0], fcore]), dipoles=dip, nbft=nbft)
    return dip


def main():
    if not os.path.exists(HARNESS):
        sys.exit("build the harness first: sh oracle/sifs/build.sh")
    rng = np.random.default_rng(2024)

    # A: 6 orbitals, 8-bit labels, 1-e arrays over several records (n1max 4), overlap and
    # kinetic / potential split, frozen-core values on every record (only the last record's
    # value of each array counts), 2-e values over several records (n2max 24)
    n = 6
    pairs = [(i, j) for i in range(1, n + 1) for j in range(1, i + 1)]
    S = [(i, j, 1.0 if i == j else 0.01 * rng.standard_normal()) for i, j in pairs]
    T = [(i, j, rng.standard_normal()) for i, j in pairs if (i + j) % 2 == 0]
    V = [(i, j, rng.standard_normal()) for i, j in pairs if (i + j) % 2 == 1]
    quads = [(i, j, k, l) for i in range(1, n + 1) for j in range(1, i + 1)
             for k in range(1, i + 1) for l in range(1, k + 1)
             if ind2(i, j) >= ind2(k, l)]
    two = [(i, j, k, l, rng.standard_normal()) for i, j, k, l in quads]
    write_case("a_fmt8", n, [n], 0, 32, 4, 64, 24, [9.25, -1.5], [-1, -2],
               [(0, 0, 0.125, S), (0, 1, -2.5, T), (0, 2, 1.75, V)], two)

    # B: the 14-orbital synthetic integrals (seed 7), 10-bit labels; H1 as kinetic
    # (diagonal) + potential (off-diagonal) so each element has one contribution; 2-e in
    # records of 64 values with a final record of a length not divisible by 8
    import jayci_amd as J
    n = 14
    h1, eri = J.synth_integrals(n, 7)
    pairs = [(i, j) for i in range(1, n + 1) for j in range(1, i + 1)]
    T = [(i, j, h1[ind2(i, j) - 1]) for i, j in pairs if i == j]
    V = [(i, j, h1[ind2(i, j) - 1]) for i, j in pairs if i != j]
    two = []
    for i in range(1, n + 1):
        for j in range(1, i + 1):
            for k in range(1, i + 1):
                for l in range(1, k + 1):
                    if ind2(i, j) >= ind2(k, l):
                        two.append((i, j, k, l, eri[ind2(ind2(i, j), ind2(k, l)) - 1]))
    assert len(two) % 64 != 0 and (len(two) % 64) % 8 != 0
    write_case("b_fmt10_synth14", n, [n], 1, 90, 64, 120, 64, [10.0], [-1],
               [(0, 1, 0.0, T), (0, 2, 0.0, V)], two)

    # C: two symmetry blocks (4 + 3 functions): H1 symmetry-blocked packed; 1-e labels
    # only inside the blocks
    nb = [4, 3]
    blocks = [(1, 4), (5, 7)]
    Hv = []
    for lo, hi in blocks:
        for i in range(lo, hi + 1):
            for j in range(lo, i + 1):
                Hv.append((i, j, rng.standard_normal()))
    two = [(i, j, k, l, rng.standard_normal()) for i, j, k, l in
           [(1, 1, 1, 1), (5, 5, 1, 1), (7, 6, 2, 1), (7, 7, 7, 7), (4, 3, 4, 3)]]
    write_case("c_sym2", 7, nb, 0, 40, 8, 40, 6, [3.0, 0.5], [-1, 0],
               [(0, 6, 0.25, Hv)], two)

    # D: dipole arrays (itypea 1, itypeb 0/1/2), each one record with last = 1, then an
    # overlap array that closes the 1-e records
    n = 5
    pairs = [(i, j) for i in range(1, n + 1) for j in range(1, i + 1)]
    comp = [[(i, j, rng.standard_normal()) for i, j in pairs] for _ in range(3)]
    S = [(i, i, 1.0) for i in range(1, n + 1)]
    write_case("d_dipole", n, [n], 0, 48, 16, 16, 4, [7.5], [-1],
               [(1, 0, 0.0, comp[0]), (1, 1, 0.0, comp[1]), (1, 2, 0.0, comp[2]), (0, 0, 0.0, S)],
               [(1, 1, 1, 1, 0.5)])
    m1 = n * (n + 1) // 2
    dip = np.zeros(3 * m1)
    for c in range(3):
        for i, j, v in comp[c]:
            dip[c * m1 + ind2(i, j) - 1] += v
    d = dict(np.load(os.path.join(OUT, "d_dipole.npz")))
    d["dipoles"] = dip
    np.savez(os.path.join(OUT, "d_dipole.npz"), **d)
    for f in sorted(os.listdir(OUT)):
        print(f, os.path.getsize(os.path.join(OUT, f)))


if __name__ == "__main__":
    main()
