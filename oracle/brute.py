"""Brute-force second quantization: H in a determinant basis from ladder operators.

TEST INFRASTRUCTURE ONLY (part of the oracle).  Independent of the Slater-Condon rules:
every element comes from applying
    H = sum_{pq,s} h_pq a+_ps a_qs + 1/2 sum_{pqrs,s,t} (pq|rs) a+_ps a+_rt a_st a_qs
to occupation-number vectors.  Spin-orbital order: alpha orbitals 0..n-1, then beta
orbitals n..2n-1; a determinant is the product of creators in ascending spin-orbital
order acting on the vacuum (alpha before beta).  Used only on tiny spaces.
"""
import numpy as np


def _annihilate(state, q):
    if not (state >> q) & 1:
        return None, 0
    sign = -1 if bin(state & ((1 << q) - 1)).count("1") & 1 else 1
    return state & ~(1 << q), sign


def _create(state, p):
    if (state >> p) & 1:
        return None, 0
    sign = -1 if bin(state & ((1 << p) - 1)).count("1") & 1 else 1
    return state | (1 << p), sign


def unpack_integrals(m1, m2, n):
    """Dense h[n,n] and g[n,n,n,n] from the 1-based packed lower-triangle layout
    (index1e(i,j) = i(i-1)/2 + j for i >= j; index2e = index1e(index1e(i,j), index1e(k,l)))."""
    def i1(i, j):
        return i * (i - 1) // 2 + j if i >= j else j * (j - 1) // 2 + i

    h = np.zeros((n, n))
    g = np.zeros((n, n, n, n))
    for i in range(1, n + 1):
        for j in range(1, n + 1):
            h[i - 1, j - 1] = m1[i1(i, j) - 1]
    for i in range(1, n + 1):
        for j in range(1, n + 1):
            for k in range(1, n + 1):
                for l in range(1, n + 1):
                    g[i - 1, j - 1, k - 1, l - 1] = m2[i1(i1(i, j), i1(k, l)) - 1]
    return h, g


def hamiltonian(dets_a, dets_b, n, h, g):
    """Dense H over the given determinants (alpha/beta masks, bit o-1 = orbital o)."""
    N = len(dets_a)
    key = {}
    for k in range(N):
        key[int(dets_a[k]) | (int(dets_b[k]) << n)] = k
    H = np.zeros((N, N))
    for k in range(N):
        st = int(dets_a[k]) | (int(dets_b[k]) << n)
        occ = [o for o in range(2 * n) if (st >> o) & 1]
        # one-body
        for q in occ:
            s1, sg1 = _annihilate(st, q)
            base = 0 if q < n else n
            for p in range(base, base + n):
                s2, sg2 = _create(s1, p)
                if s2 is None or s2 not in key:
                    continue
                H[key[s2], k] += sg1 * sg2 * h[p - base, q - base]
        # two-body: 1/2 (pq|rs) a+_p a+_r a_s a_q, p~q same spin, r~s same spin
        for q in occ:
            s1, sg1 = _annihilate(st, q)
            bq = 0 if q < n else n
            for s in occ:
                if s == q:
                    continue
                s2, sg2 = _annihilate(s1, s)
                bs = 0 if s < n else n
                for r in range(bs, bs + n):
                    s3, sg3 = _create(s2, r)
                    if s3 is None:
                        continue
                    for p in range(bq, bq + n):
                        s4, sg4 = _create(s3, p)
                        if s4 is None or s4 not in key:
                            continue
                        H[key[s4], k] += 0.5 * sg1 * sg2 * sg3 * sg4 * \
                            g[p - bq, q - bq, r - bs, s - bs]
    return H
