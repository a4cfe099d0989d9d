// Slater-Condon matrix elements on 64-bit orbital masks, shared by host and device code.
//
// Rules follow mpi_source/action_util.c:167-258 (hmatels_0x / 1x / 2xaa / 2xab), the
// corrected same-spin double of the reference (the serial source/action_util.c:511-512
// pairing is wrong).  Orbital o (1-based) is bit o-1.  Integrals use the packed 8-fold
// layout of source/moindex.c:14-41; with 0-based orbitals the flat index of (ij|kl) is
// tri(tri(i,j), tri(k,l)) where tri(a,b) = max(a,b)*(max(a,b)+1)/2 + min(a,b).
//
// Phase conventions:
//   single i->f in S:       (-1)^popcount(S & bits strictly between i and f)
//   same-spin double:       i0<i1 -> f0<f1, sign = s(i0->f0 in I) * s(i1->f1 in I-i0+f0),
//                           value = sign * [(i0 f0|i1 f1) - (i0 f1|i1 f0)]
//   alpha-beta double:      sign = s_alpha * s_beta, value = sign * (ia fa|ib fb)
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define JCI_HD __host__ __device__ __forceinline__
#else
#define JCI_HD static inline
#endif

JCI_HD int jci_popc(uint64_t x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __popcll(x);
#else
    return __builtin_popcountll(x);
#endif
}

JCI_HD int jci_ctz(uint64_t x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __ffsll((unsigned long long)x) - 1;
#else
    return __builtin_ctzll(x);
#endif
}

JCI_HD long long jci_tri(long long a, long long b)
{
    return a >= b ? a * (a + 1) / 2 + b : b * (b + 1) / 2 + a;
}

// (ij|kl), 0-based orbitals
JCI_HD double jci_eri(const double *m2, int i, int j, int k, int l)
{
    return m2[jci_tri(jci_tri(i, j), jci_tri(k, l))];
}

JCI_HD double jci_h1(const double *m1, int i, int j) { return m1[jci_tri(i, j)]; }

// mask of orbitals strictly between 0-based positions a and b (a != b)
JCI_HD uint64_t jci_between(int a, int b)
{
    int lo = a < b ? a : b, hi = a < b ? b : a;
    return ((1ULL << hi) - 1ULL) & ~((2ULL << lo) - 1ULL);
}

JCI_HD int jci_sign(uint64_t s, int i, int f)
{
    return (jci_popc(s & jci_between(i, f)) & 1) ? -1 : 1;
}

// <I|H|I>
JCI_HD double jci_diag(uint64_t a, uint64_t b, const double *m1, const double *m2)
{
    double v = 0.0;
    for (uint64_t x = a; x; x &= x - 1) {
        int i = jci_ctz(x);
        v += jci_h1(m1, i, i);
        for (uint64_t y = a & ((1ULL << i) - 1ULL); y; y &= y - 1) {
            int j = jci_ctz(y);
            v += jci_eri(m2, i, i, j, j) - jci_eri(m2, i, j, i, j);
        }
        for (uint64_t y = b; y; y &= y - 1) {
            int j = jci_ctz(y);
            v += jci_eri(m2, i, i, j, j);
        }
    }
    for (uint64_t x = b; x; x &= x - 1) {
        int i = jci_ctz(x);
        v += jci_h1(m1, i, i);
        for (uint64_t y = b & ((1ULL << i) - 1ULL); y; y &= y - 1) {
            int j = jci_ctz(y);
            v += jci_eri(m2, i, i, j, j) - jci_eri(m2, i, j, i, j);
        }
    }
    return v;
}

// single i->f in string s1 (the excited spin of the row determinant), s2 the other spin;
// returns the unsigned bracket (multiply by jci_sign(s1, i, f)).
JCI_HD double jci_single_bracket(uint64_t s1, uint64_t s2, int i, int f,
                                 const double *m1, const double *m2)
{
    double v = jci_h1(m1, i, f);
    for (uint64_t x = s1 & ~(1ULL << i); x; x &= x - 1) {
        int k = jci_ctz(x);
        v += jci_eri(m2, k, k, i, f) - jci_eri(m2, k, i, k, f);
    }
    for (uint64_t x = s2; x; x &= x - 1) {
        int k = jci_ctz(x);
        v += jci_eri(m2, k, k, i, f);
    }
    return v;
}

// same-spin double (i0<i1 removed from s, f0<f1 added); returns value, sign in *ph
JCI_HD double jci_double_same(uint64_t s, int i0, int i1, int f0, int f1, const double *m2,
                              int *ph)
{
    int s1 = jci_sign(s, i0, f0);
    uint64_t k = (s & ~(1ULL << i0)) | (1ULL << f0);
    int sg = s1 * jci_sign(k, i1, f1);
    *ph = sg;
    return sg * (jci_eri(m2, i0, f0, i1, f1) - jci_eri(m2, i0, f1, i1, f0));
}

// General <I|H|J> from masks; *deg = excitation degree, *phase the sign (0 if disconnected).
JCI_HD double jci_hij(uint64_t ia, uint64_t ib, uint64_t ja, uint64_t jb,
                      const double *m1, const double *m2, int *deg, int *phase)
{
    uint64_t xa = ia ^ ja, xb = ib ^ jb;
    int na = jci_popc(xa & ia), nb = jci_popc(xb & ib);
    int nfa = jci_popc(xa & ja), nfb = jci_popc(xb & jb);
    *deg = na + nb;
    *phase = 0;
    if (na != nfa || nb != nfb || na + nb > 2) return 0.0;
    if (na + nb == 0) {
        *phase = 1;
        return jci_diag(ia, ib, m1, m2);
    }
    if (na + nb == 1) {
        uint64_t s1 = na ? ia : ib, s2 = na ? ib : ia, x = na ? xa : xb;
        int i = jci_ctz(x & s1), f = jci_ctz(x & ~s1);
        int sg = jci_sign(s1, i, f);
        *phase = sg;
        return sg * jci_single_bracket(s1, s2, i, f, m1, m2);
    }
    if (na == 1) {
        int i0 = jci_ctz(xa & ia), f0 = jci_ctz(xa & ja);
        int i1 = jci_ctz(xb & ib), f1 = jci_ctz(xb & jb);
        int sg = jci_sign(ia, i0, f0) * jci_sign(ib, i1, f1);
        *phase = sg;
        return sg * jci_eri(m2, i0, f0, i1, f1);
    }
    uint64_t s = na == 2 ? ia : ib, x = na == 2 ? xa : xb, t = na == 2 ? ja : jb;
    uint64_t io = x & s, fo = x & t;
    int i0 = jci_ctz(io), i1 = jci_ctz(io & (io - 1));
    int f0 = jci_ctz(fo), f1 = jci_ctz(fo & (fo - 1));
    return jci_double_same(s, i0, i1, f0, f1, m2, phase);
}

// ---- string addressing -------------------------------------------------------------
// Lexicographic rank of a k-subset of an m-orbital subspace (bits relative to the
// subspace start), via the complement/colex identity
//   rank = C(m,k) - 1 - sum_t C(m - c_t, k + 1 - t),  c_1 < ... < c_k (1-based),
// which is the order generated by source/straddress.c:77 (str_strfind1).
JCI_HD long long jci_lex_rank(uint64_t x, int m, int k, const unsigned long long *binom)
{
    long long r = (long long)binom[m * 65 + k] - 1;
    int t = 1;
    for (; x; x &= x - 1, t++) {
        int c = jci_ctz(x) + 1;
        r -= (long long)binom[(m - c) * 65 + (k + 1 - t)];
    }
    return r;
}

// inverse of jci_lex_rank: mask (relative to subspace start) of lexicographic rank r
JCI_HD uint64_t jci_lex_unrank(long long r, int m, int k, const unsigned long long *binom)
{
    uint64_t mask = 0;
    int c = 0;
    for (int t = 1; t <= k; t++) {
        c++;
        for (;;) {
            long long cnt = (long long)binom[(m - c) * 65 + (k - t)];
            if (r < cnt) break;
            r -= cnt;
            c++;
        }
        mask |= 1ULL << (c - 1);
    }
    return mask;
}
