/*
 * jayci_oracle.c -- CPU restatement of the JayCI sigma (Hv) / Davidson path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the oracle: only tests/, the smoke check and
 * bench.py's cpu_baseline leg load it (as oracle/liboracle.so through ctypes), and
 * only as the checker.  The product library never links, loads or calls it.
 *
 * Every routine follows a reference routine (paths relative to the reference tree):
 *   orc_groups ........... source/citruncate.c:205-251   (allocate_eospace_array)
 *   string order ......... source/citruncate.c:321-415   (compute_eostrings),
 *                          source/citruncate.c:543-575   (generate_string_list),
 *                          source/straddress.c:77-100    (str_strfind1, lexicographic successor)
 *   determinant blocks ... source/citruncate.c:291-312   (serial rule, compute_detnum),
 *                          mpi_source/citruncate.c:414-471 (MRCI-SD rule, compute_detnum),
 *                          source/citruncate.c:497-538   (block order, alpha-major inside a block)
 *   integral addressing .. source/moindex.c:14-41        (index1e / index2e, 1-based packed)
 *   matrix elements ...... mpi_source/action_util.c:167-258 (hmatels_0x / 1x / 2xaa / 2xab)
 *   sigma ................ source/action_util.c:121-155  (interface: c += H v),
 *                          mpi_source/pdavidson.c:1964-2067 (string-driven rows, owner computes)
 *   Davidson ............. source/davidson.c:91-447, source/mathutil.c:169-204 (MGS),
 *                          source/initguess_sbd.c:24-109 (leading-block guess)
 *
 * Conventions (shared with the product, stated here so the oracle stands alone):
 *   - orbitals are 1-based; orbital o is bit (o-1) of a 64-bit string mask;
 *   - a determinant is |alpha string> x |beta string>, creators in ascending orbital
 *     order, all alpha creators left of all beta creators;
 *   - single excitation i->f in string S: sign = (-1)^(number of electrons of S on
 *     orbitals strictly between i and f);
 *   - same-spin double (i0<i1 -> f0<f1): sign = s(i0->f0 in I) * s(i1->f1 in K) with
 *     K = I - i0 + f0, and <I|H|J> = sign * [(i0 f0|i1 f1) - (i0 f1|i1 f0)];
 *   - alpha-beta double: sign = s_alpha * s_beta, <I|H|J> = sign * (ia fa|ib fb).
 *   The "phase" reported for a pair is that sign (+1 for the diagonal).
 *
 * Deliberately written with plain loops over orbitals (no popcount/ffs tricks, no
 * combinatorial ranking), so that it shares no arithmetic shortcuts with the GPU path:
 * string addresses come from binary search over the enumerated strings.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_MAXG 256

typedef struct {
    int nelec, norbs, ndocc, nactv, xlvl;
    int ngrp;
    int gd[ORC_MAXG], ga[ORC_MAXG], gv[ORC_MAXG];
    long long gstart[ORC_MAXG], gn[ORC_MAXG];
    long long nstr;
    uint64_t *mask;      /* reference order */
    uint64_t *smask;     /* sorted copy for address lookup */
    long long *sidx;     /* index of smask[k] in reference order */
} orc_strs;

typedef struct orc_space {
    int na, nb, norbs, ndocc, nactv, xlvl, rule;
    orc_strs s[2];
    int nblk;
    int *bp, *bq;
    long long *boff;     /* nblk+1 */
    int *blkid;          /* [ngrp_a * ngrp_b] -> block index or -1 */
    long long ndets;
} orc_space;

static int imin(int a, int b) { return a < b ? a : b; }
static int imax(int a, int b) { return a > b ? a : b; }

static long long binom(int n, int k)
{
    if (k < 0 || n < 0 || k > n) return 0;
    long long r = 1;
    for (int i = 1; i <= k; i++) r = r * (n - k + i) / i;
    return r;
}

/* ---- groups: source/citruncate.c:205-251 ------------------------------------ */
int orc_groups(int nelec, int norbs, int ndocc, int nactv, int xlvl,
               int *docc, int *actv, int *virt, long long *nstr, int maxg)
{
    int vorbs = norbs - ndocc - nactv;
    int min_d = ndocc - imin(ndocc, xlvl);
    int max_d = ndocc;
    int min_a = (nelec - ndocc) - imin(nelec - ndocc, xlvl);
    int max_a = imin(imin(ndocc, xlvl) + (nelec - ndocc), nactv);
    int n = 0;
    for (int i = max_d; i >= min_d; i--)
        for (int j = max_a; j >= min_a; j--)
            for (int k = 0; k <= imin(vorbs, xlvl); k++) {
                if (i + j + k != nelec) continue;
                if (n < maxg) {
                    docc[n] = i; actv[n] = j; virt[n] = k;
                    nstr[n] = binom(ndocc, i) * binom(nactv, j) * binom(vorbs, k);
                }
                n++;
            }
    return n;
}

/* lexicographic successor of a k-combination of {1..m}; returns 0 at the end.
 * Same order as source/straddress.c:77 (str_strfind1). */
static int lex_next(int *c, int k, int m)
{
    int t = k - 1;
    while (t >= 0 && c[t] == m - (k - 1 - t)) t--;
    if (t < 0) return 0;
    c[t]++;
    for (int u = t + 1; u < k; u++) c[u] = c[u - 1] + 1;
    return 1;
}

static uint64_t combo_mask(const int *c, int k, int off)
{
    uint64_t m = 0;
    for (int i = 0; i < k; i++) m |= 1ULL << (c[i] + off - 1);
    return m;
}

static int cmp_u64(const void *a, const void *b)
{
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

/* strings of one spin, in reference order (groups, then docc / actv / virt nesting) */
static int strs_build(orc_strs *s, int nelec, int norbs, int ndocc, int nactv, int xlvl)
{
    memset(s, 0, sizeof(*s));
    s->nelec = nelec; s->norbs = norbs; s->ndocc = ndocc; s->nactv = nactv; s->xlvl = xlvl;
    s->ngrp = orc_groups(nelec, norbs, ndocc, nactv, xlvl, s->gd, s->ga, s->gv, s->gn, ORC_MAXG);
    if (s->ngrp > ORC_MAXG) return -1;
    long long tot = 0;
    for (int g = 0; g < s->ngrp; g++) { s->gstart[g] = tot; tot += s->gn[g]; }
    s->nstr = tot;
    s->mask = (uint64_t *)malloc(sizeof(uint64_t) * (tot ? tot : 1));
    int vorbs = norbs - ndocc - nactv;
    long long p = 0;
    for (int g = 0; g < s->ngrp; g++) {
        int kd = s->gd[g], ka = s->ga[g], kv = s->gv[g];
        int cd[64], ca[64], cv[64];
        for (int i = 0; i < kd; i++) cd[i] = i + 1;
        do {
            for (int i = 0; i < ka; i++) ca[i] = i + 1;
            do {
                for (int i = 0; i < kv; i++) cv[i] = i + 1;
                do {
                    s->mask[p++] = combo_mask(cd, kd, 0) | combo_mask(ca, ka, ndocc) |
                                   combo_mask(cv, kv, ndocc + nactv);
                } while (kv > 0 && lex_next(cv, kv, vorbs));
            } while (ka > 0 && lex_next(ca, ka, nactv));
        } while (kd > 0 && lex_next(cd, kd, ndocc));
    }
    if (p != tot) return -2;
    /* sorted lookup table */
    s->smask = (uint64_t *)malloc(sizeof(uint64_t) * (tot ? tot : 1));
    s->sidx = (long long *)malloc(sizeof(long long) * (tot ? tot : 1));
    uint64_t *pairs = (uint64_t *)malloc(sizeof(uint64_t) * 2 * (tot ? tot : 1));
    for (long long i = 0; i < tot; i++) { pairs[2 * i] = s->mask[i]; pairs[2 * i + 1] = (uint64_t)i; }
    qsort(pairs, (size_t)tot, 2 * sizeof(uint64_t), cmp_u64);
    for (long long i = 0; i < tot; i++) { s->smask[i] = pairs[2 * i]; s->sidx[i] = (long long)pairs[2 * i + 1]; }
    free(pairs);
    return 0;
}

static void strs_free(orc_strs *s)
{
    free(s->mask); free(s->smask); free(s->sidx);
}

static long long strs_find(const orc_strs *s, uint64_t m)
{
    long long lo = 0, hi = s->nstr - 1;
    while (lo <= hi) {
        long long mid = (lo + hi) / 2;
        if (s->smask[mid] == m) return s->sidx[mid];
        if (s->smask[mid] < m) lo = mid + 1; else hi = mid - 1;
    }
    return -1;
}

static int strs_group(const orc_strs *s, long long idx)
{
    int lo = 0, hi = s->ngrp - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) / 2;
        if (s->gstart[mid] <= idx) lo = mid; else hi = mid - 1;
    }
    return lo;
}

/* ---- space ------------------------------------------------------------------ */
orc_space *orc_space_new(int na, int nb, int norbs, int ndocc, int nactv, int xlvl, int rule)
{
    if (norbs > 64 || norbs < 1) return NULL;
    orc_space *sp = (orc_space *)calloc(1, sizeof(orc_space));
    sp->na = na; sp->nb = nb; sp->norbs = norbs; sp->ndocc = ndocc; sp->nactv = nactv;
    sp->xlvl = xlvl; sp->rule = rule;
    if (strs_build(&sp->s[0], na, norbs, ndocc, nactv, xlvl) ||
        strs_build(&sp->s[1], nb, norbs, ndocc, nactv, xlvl)) { free(sp); return NULL; }
    int ga = sp->s[0].ngrp, gb = sp->s[1].ngrp;
    sp->bp = (int *)malloc(sizeof(int) * (ga * gb + 1));
    sp->bq = (int *)malloc(sizeof(int) * (ga * gb + 1));
    sp->boff = (long long *)malloc(sizeof(long long) * (ga * gb + 1));
    sp->blkid = (int *)malloc(sizeof(int) * (ga * gb + 1));
    /* serial rule: source/citruncate.c:300-307; MRCI-SD rule: mpi_source/citruncate.c:427-439 */
    int doccmin, actvmin;
    if (rule == 0) { doccmin = 2 * (ndocc - imin(ndocc, xlvl)); actvmin = -1; }
    else {
        doccmin = imax(2 * ndocc - xlvl, 0);
        actvmin = imax((na - ndocc) + (nb - ndocc) - xlvl, 0);
    }
    long long off = 0;
    int nb_ = 0;
    for (int p = 0; p < ga; p++)
        for (int q = 0; q < gb; q++) {
            sp->blkid[p * gb + q] = -1;
            if (sp->s[0].gd[p] + sp->s[1].gd[q] < doccmin) continue;
            if (rule != 0 && sp->s[0].ga[p] + sp->s[1].ga[q] < actvmin) continue;
            if (sp->s[0].gv[p] + sp->s[1].gv[q] > xlvl) continue;
            sp->bp[nb_] = p; sp->bq[nb_] = q; sp->boff[nb_] = off;
            sp->blkid[p * gb + q] = nb_;
            off += sp->s[0].gn[p] * sp->s[1].gn[q];
            nb_++;
        }
    sp->boff[nb_] = off;
    sp->nblk = nb_;
    sp->ndets = off;
    return sp;
}

void orc_space_free(orc_space *sp)
{
    if (!sp) return;
    strs_free(&sp->s[0]); strs_free(&sp->s[1]);
    free(sp->bp); free(sp->bq); free(sp->boff); free(sp->blkid);
    free(sp);
}

long long orc_space_ndets(const orc_space *sp) { return sp->ndets; }
long long orc_space_nstr(const orc_space *sp, int spin) { return sp->s[spin].nstr; }
int orc_space_ngroups(const orc_space *sp, int spin) { return sp->s[spin].ngrp; }
int orc_space_nblocks(const orc_space *sp) { return sp->nblk; }

void orc_space_strings(const orc_space *sp, int spin, uint64_t *out)
{
    memcpy(out, sp->s[spin].mask, sizeof(uint64_t) * sp->s[spin].nstr);
}

/* groups as rows of (docc, actv, virt, start, nstr) */
void orc_space_groups(const orc_space *sp, int spin, long long *out)
{
    const orc_strs *s = &sp->s[spin];
    for (int g = 0; g < s->ngrp; g++) {
        out[5 * g] = s->gd[g]; out[5 * g + 1] = s->ga[g]; out[5 * g + 2] = s->gv[g];
        out[5 * g + 3] = s->gstart[g]; out[5 * g + 4] = s->gn[g];
    }
}

/* blocks as rows of (p, q, offset) */
void orc_space_blocks(const orc_space *sp, long long *out)
{
    for (int b = 0; b < sp->nblk; b++) {
        out[3 * b] = sp->bp[b]; out[3 * b + 1] = sp->bq[b]; out[3 * b + 2] = sp->boff[b];
    }
}

/* determinant index of (alpha string index, beta string index), or -1 */
static long long det_index(const orc_space *sp, long long ia, long long ib)
{
    int p = strs_group(&sp->s[0], ia), q = strs_group(&sp->s[1], ib);
    int b = sp->blkid[p * sp->s[1].ngrp + q];
    if (b < 0) return -1;
    return sp->boff[b] + (ia - sp->s[0].gstart[p]) * sp->s[1].gn[q] + (ib - sp->s[1].gstart[q]);
}

long long orc_space_index(const orc_space *sp, uint64_t ma, uint64_t mb)
{
    long long ia = strs_find(&sp->s[0], ma), ib = strs_find(&sp->s[1], mb);
    if (ia < 0 || ib < 0) return -1;
    return det_index(sp, ia, ib);
}

/* all determinants as (alpha mask, beta mask), reference order */
void orc_space_dets(const orc_space *sp, uint64_t *a, uint64_t *b)
{
    long long k = 0;
    for (int blk = 0; blk < sp->nblk; blk++) {
        int p = sp->bp[blk], q = sp->bq[blk];
        for (long long i = 0; i < sp->s[0].gn[p]; i++)
            for (long long j = 0; j < sp->s[1].gn[q]; j++) {
                a[k] = sp->s[0].mask[sp->s[0].gstart[p] + i];
                b[k] = sp->s[1].mask[sp->s[1].gstart[q] + j];
                k++;
            }
    }
}

void orc_space_det(const orc_space *sp, long long idx, uint64_t *a, uint64_t *b)
{
    int lo = 0, hi = sp->nblk - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) / 2;
        if (sp->boff[mid] <= idx) lo = mid; else hi = mid - 1;
    }
    int p = sp->bp[lo], q = sp->bq[lo];
    long long r = idx - sp->boff[lo];
    long long nq = sp->s[1].gn[q];
    *a = sp->s[0].mask[sp->s[0].gstart[p] + r / nq];
    *b = sp->s[1].mask[sp->s[1].gstart[q] + r % nq];
}

/* ---- integrals: source/moindex.c:14-41 --------------------------------------- */
static inline long ix1(long i, long j) { return i > j ? (i * (i - 1)) / 2 + j : (j * (j - 1)) / 2 + i; }
static inline long ix2(long i, long j, long k, long l) { return ix1(ix1(i, j), ix1(k, l)); }
#define ONE(i, j) m1[ix1((i), (j)) - 1]
#define TWO(i, j, k, l) m2[ix2((i), (j), (k), (l)) - 1]

long orc_index1e(long i, long j) { return ix1(i, j); }
long orc_index2e(long i, long j, long k, long l) { return ix2(i, j, k, l); }

/* ---- Slater-Condon: mpi_source/action_util.c:167-258 -------------------------- */
static int occ_list(uint64_t m, int norbs, int *o)
{
    int k = 0;
    for (int i = 1; i <= norbs; i++) if ((m >> (i - 1)) & 1ULL) o[k++] = i;
    return k;
}

static int exc_sign(uint64_t m, int i, int f)
{
    int lo = i < f ? i : f, hi = i < f ? f : i, cnt = 0;
    for (int o = lo + 1; o < hi; o++) if ((m >> (o - 1)) & 1ULL) cnt++;
    return (cnt & 1) ? -1 : 1;
}

static double diag_elem(uint64_t ia, uint64_t ib, int norbs, const double *m1, const double *m2)
{
    int oa[64], ob[64];
    int na = occ_list(ia, norbs, oa), nb = occ_list(ib, norbs, ob);
    double v = 0.0;
    for (int i = 0; i < na; i++) {
        v += ONE(oa[i], oa[i]);
        for (int j = 0; j < i; j++) v += TWO(oa[i], oa[i], oa[j], oa[j]) - TWO(oa[i], oa[j], oa[i], oa[j]);
    }
    for (int i = 0; i < nb; i++) {
        v += ONE(ob[i], ob[i]);
        for (int j = 0; j < i; j++) v += TWO(ob[i], ob[i], ob[j], ob[j]) - TWO(ob[i], ob[j], ob[i], ob[j]);
    }
    for (int i = 0; i < na; i++)
        for (int j = 0; j < nb; j++) v += TWO(oa[i], oa[i], ob[j], ob[j]);
    return v;
}

/* single i->f in string `s1` (the row's string of the excited spin), `s2` the other string */
static double single_elem(uint64_t s1, uint64_t s2, int i, int f, int norbs,
                          const double *m1, const double *m2, int *phase)
{
    int o1[64], o2[64];
    int n1 = occ_list(s1, norbs, o1), n2 = occ_list(s2, norbs, o2);
    int s = exc_sign(s1, i, f);
    double v = ONE(i, f);
    for (int k = 0; k < n1; k++)
        if (o1[k] != i) v += TWO(o1[k], o1[k], i, f) - TWO(o1[k], i, o1[k], f);
    for (int k = 0; k < n2; k++) v += TWO(o2[k], o2[k], i, f);
    *phase = s;
    return s * v;
}

static int diff_orbs(uint64_t from, uint64_t to, int norbs, int *out)
{
    /* orbitals in `from` but not in `to`, ascending */
    int k = 0;
    for (int o = 1; o <= norbs; o++)
        if (((from >> (o - 1)) & 1ULL) && !((to >> (o - 1)) & 1ULL)) out[k++] = o;
    return k;
}

double orc_hij(uint64_t ia, uint64_t ib, uint64_t ja, uint64_t jb, int norbs,
               const double *m1, const double *m2, int *deg, int *phase)
{
    int ioa[64], foa[64], iob[64], fob[64];
    int na = diff_orbs(ia, ja, norbs, ioa), nfa = diff_orbs(ja, ia, norbs, foa);
    int nb = diff_orbs(ib, jb, norbs, iob), nfb = diff_orbs(jb, ib, norbs, fob);
    int d = na + nb;
    *deg = d;
    *phase = 0;
    if (na != nfa || nb != nfb || d > 2) return 0.0;
    if (d == 0) { *phase = 1; return diag_elem(ia, ib, norbs, m1, m2); }
    if (d == 1) {
        if (na == 1) return single_elem(ia, ib, ioa[0], foa[0], norbs, m1, m2, phase);
        return single_elem(ib, ia, iob[0], fob[0], norbs, m1, m2, phase);
    }
    if (na == 1) {   /* alpha-beta double: hmatels_2xab */
        int s = exc_sign(ia, ioa[0], foa[0]) * exc_sign(ib, iob[0], fob[0]);
        *phase = s;
        return s * TWO(ioa[0], foa[0], iob[0], fob[0]);
    }
    /* same-spin double: hmatels_2xaa */
    uint64_t str = na == 2 ? ia : ib;
    const int *io = na == 2 ? ioa : iob;
    const int *fo = na == 2 ? foa : fob;
    int s1 = exc_sign(str, io[0], fo[0]);
    uint64_t k = (str & ~(1ULL << (io[0] - 1))) | (1ULL << (fo[0] - 1));
    int s2 = exc_sign(k, io[1], fo[1]);
    *phase = s1 * s2;
    return s1 * s2 * (TWO(io[0], fo[0], io[1], fo[1]) - TWO(io[0], fo[1], io[1], fo[0]));
}

/* ---- string excitations ------------------------------------------------------ */
typedef struct { long long idx; int deg; } exc_t;

/* all strings of the space at excitation degree 1 or 2 from mask m (m itself excluded) */
static long long str_excitations(const orc_strs *s, uint64_t m, exc_t *out, long long cap)
{
    int occ[64], vir[64], no = 0, nv = 0;
    for (int o = 1; o <= s->norbs; o++) {
        if ((m >> (o - 1)) & 1ULL) occ[no++] = o; else vir[nv++] = o;
    }
    long long n = 0;
    for (int a = 0; a < no; a++)
        for (int b = 0; b < nv; b++) {
            uint64_t t = (m & ~(1ULL << (occ[a] - 1))) | (1ULL << (vir[b] - 1));
            long long j = strs_find(s, t);
            if (j >= 0) { if (n < cap) { out[n].idx = j; out[n].deg = 1; } n++; }
        }
    for (int a0 = 0; a0 < no; a0++)
        for (int a1 = a0 + 1; a1 < no; a1++)
            for (int b0 = 0; b0 < nv; b0++)
                for (int b1 = b0 + 1; b1 < nv; b1++) {
                    uint64_t t = m & ~(1ULL << (occ[a0] - 1)) & ~(1ULL << (occ[a1] - 1));
                    t |= (1ULL << (vir[b0] - 1)) | (1ULL << (vir[b1] - 1));
                    long long j = strs_find(s, t);
                    if (j >= 0) { if (n < cap) { out[n].idx = j; out[n].deg = 2; } n++; }
                }
    return n;
}

static long long max_exc(const orc_strs *s)
{
    long long no = s->nelec, nv = s->norbs - s->nelec;
    return no * nv + binom((int)no, 2) * binom((int)nv, 2) + 1;
}

typedef struct { long long col; double val; int phase, deg; } ent_t;

static int cmp_ent(const void *a, const void *b)
{
    long long x = ((const ent_t *)a)->col, y = ((const ent_t *)b)->col;
    return x < y ? -1 : (x > y ? 1 : 0);
}

/* connected off-diagonal entries of row I, both triangles, unsorted.  la/lb are the
 * excitation lists of I's alpha and beta strings. */
static long long row_entries(const orc_space *sp, long long ia, long long ib,
                             const exc_t *la, long long nla, const exc_t *lb, long long nlb,
                             const double *m1, const double *m2, ent_t *out)
{
    const orc_strs *A = &sp->s[0], *B = &sp->s[1];
    uint64_t ma = A->mask[ia], mb = B->mask[ib];
    long long n = 0;
    /* beta-only excitations */
    for (long long y = 0; y < nlb; y++) {
        long long J = det_index(sp, ia, lb[y].idx);
        if (J < 0) continue;
        int d, ph;
        double v = orc_hij(ma, mb, ma, B->mask[lb[y].idx], sp->norbs, m1, m2, &d, &ph);
        out[n].col = J; out[n].val = v; out[n].phase = ph; out[n].deg = d; n++;
    }
    for (long long x = 0; x < nla; x++) {
        long long ja = la[x].idx;
        /* alpha-only */
        long long J = det_index(sp, ja, ib);
        if (J >= 0) {
            int d, ph;
            double v = orc_hij(ma, mb, A->mask[ja], mb, sp->norbs, m1, m2, &d, &ph);
            out[n].col = J; out[n].val = v; out[n].phase = ph; out[n].deg = d; n++;
        }
        if (la[x].deg != 1) continue;
        /* alpha single x beta single */
        for (long long y = 0; y < nlb; y++) {
            if (lb[y].deg != 1) continue;
            J = det_index(sp, ja, lb[y].idx);
            if (J < 0) continue;
            int d, ph;
            double v = orc_hij(ma, mb, A->mask[ja], B->mask[lb[y].idx], sp->norbs, m1, m2, &d, &ph);
            out[n].col = J; out[n].val = v; out[n].phase = ph; out[n].deg = d; n++;
        }
    }
    return n;
}

static void det_strings(const orc_space *sp, long long I, long long *ia, long long *ib)
{
    int lo = 0, hi = sp->nblk - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) / 2;
        if (sp->boff[mid] <= I) lo = mid; else hi = mid - 1;
    }
    int p = sp->bp[lo], q = sp->bq[lo];
    long long r = I - sp->boff[lo], nq = sp->s[1].gn[q];
    *ia = sp->s[0].gstart[p] + r / nq;
    *ib = sp->s[1].gstart[q] + r % nq;
}

/* Row I of H, off-diagonal connected entries (both triangles), sorted by column.
 * Returns the count (entries written only up to cap). */
long long orc_row(const orc_space *sp, long long I, const double *m1, const double *m2,
                  long long *cols, double *vals, int *phase, int *deg, long long cap)
{
    long long ia, ib;
    det_strings(sp, I, &ia, &ib);
    long long ca = max_exc(&sp->s[0]), cb = max_exc(&sp->s[1]);
    exc_t *la = (exc_t *)malloc(sizeof(exc_t) * ca), *lb = (exc_t *)malloc(sizeof(exc_t) * cb);
    long long nla = str_excitations(&sp->s[0], sp->s[0].mask[ia], la, ca);
    long long nlb = str_excitations(&sp->s[1], sp->s[1].mask[ib], lb, cb);
    long long nsingle_a = 0, nsingle_b = 0;
    for (long long x = 0; x < nla; x++) nsingle_a += la[x].deg == 1;
    for (long long y = 0; y < nlb; y++) nsingle_b += lb[y].deg == 1;
    long long bound = nla + nlb + nsingle_a * nsingle_b + 1;
    ent_t *e = (ent_t *)malloc(sizeof(ent_t) * bound);
    long long n = row_entries(sp, ia, ib, la, nla, lb, nlb, m1, m2, e);
    qsort(e, (size_t)n, sizeof(ent_t), cmp_ent);
    for (long long k = 0; k < n && k < cap; k++) {
        if (cols) cols[k] = e[k].col;
        if (vals) vals[k] = e[k].val;
        if (phase) phase[k] = e[k].phase;
        if (deg) deg[k] = e[k].deg;
    }
    free(la); free(lb); free(e);
    return n;
}

/* H_II for all determinants (source/action_util.c:106, compute_hdgls) */
void orc_hdiag(const orc_space *sp, const double *m1, const double *m2, double *hd, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    #pragma omp parallel for schedule(static)
    for (long long I = 0; I < sp->ndets; I++) {
        long long ia, ib;
        det_strings(sp, I, &ia, &ib);
        hd[I] = diag_elem(sp->s[0].mask[ia], sp->s[1].mask[ib], sp->norbs, m1, m2);
    }
}

/* sigma for a list of rows: c_rows[r + k*nrows] = sum_J H[rows[r], J] v[J + k*N].
 * Owner computes the full row (no scatter), summation in ascending column order. */
void orc_sigma_rows(const orc_space *sp, const double *m1, const double *m2, int nvec,
                    const double *v, const long long *rows, long long nrows, double *c_rows,
                    int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    long long N = sp->ndets;
    long long ca = max_exc(&sp->s[0]), cb = max_exc(&sp->s[1]);
    #pragma omp parallel
    {
        exc_t *la = (exc_t *)malloc(sizeof(exc_t) * ca), *lb = (exc_t *)malloc(sizeof(exc_t) * cb);
        long long ecap = 0;
        ent_t *e = NULL;
        #pragma omp for schedule(dynamic, 16)
        for (long long r = 0; r < nrows; r++) {
            long long I = rows[r], ia, ib;
            det_strings(sp, I, &ia, &ib);
            long long nla = str_excitations(&sp->s[0], sp->s[0].mask[ia], la, ca);
            long long nlb = str_excitations(&sp->s[1], sp->s[1].mask[ib], lb, cb);
            long long s1a = 0, s1b = 0;
            for (long long x = 0; x < nla; x++) s1a += la[x].deg == 1;
            for (long long y = 0; y < nlb; y++) s1b += lb[y].deg == 1;
            long long bound = nla + nlb + s1a * s1b + 1;
            if (bound > ecap) { free(e); ecap = bound; e = (ent_t *)malloc(sizeof(ent_t) * ecap); }
            long long n = row_entries(sp, ia, ib, la, nla, lb, nlb, m1, m2, e);
            qsort(e, (size_t)n, sizeof(ent_t), cmp_ent);
            double hii = diag_elem(sp->s[0].mask[ia], sp->s[1].mask[ib], sp->norbs, m1, m2);
            for (int k = 0; k < nvec; k++) {
                const double *vk = v + (size_t)k * N;
                double acc = hii * vk[I];
                for (long long t = 0; t < n; t++) acc += e[t].val * vk[e[t].col];
                c_rows[r + (size_t)k * nrows] = acc;
            }
        }
        free(la); free(lb); free(e);
    }
}

/* ---- precomputed-list sigma for whole spaces ----------------------------------- */
typedef struct {
    long long *off;   /* nstr+1 */
    exc_t *ent;
} exc_table;

static void exc_table_build(const orc_strs *s, exc_table *t)
{
    long long cap = max_exc(s);
    t->off = (long long *)malloc(sizeof(long long) * (s->nstr + 1));
    long long *cnt = (long long *)malloc(sizeof(long long) * (s->nstr + 1));
    #pragma omp parallel
    {
        exc_t *tmp = (exc_t *)malloc(sizeof(exc_t) * cap);
        #pragma omp for schedule(dynamic, 64)
        for (long long i = 0; i < s->nstr; i++) cnt[i] = str_excitations(s, s->mask[i], tmp, cap);
        free(tmp);
    }
    t->off[0] = 0;
    for (long long i = 0; i < s->nstr; i++) t->off[i + 1] = t->off[i] + cnt[i];
    t->ent = (exc_t *)malloc(sizeof(exc_t) * (t->off[s->nstr] + 1));
    #pragma omp parallel for schedule(dynamic, 64)
    for (long long i = 0; i < s->nstr; i++) str_excitations(s, s->mask[i], t->ent + t->off[i], cnt[i]);
    free(cnt);
}

static void exc_table_free(exc_table *t) { free(t->off); free(t->ent); }

/* c = H v for all rows (overwrites c), vectors stored one after another (N each) */
void orc_sigma(const orc_space *sp, const double *m1, const double *m2, int nvec,
               const double *v, double *c, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    long long N = sp->ndets;
    exc_table ta, tb;
    exc_table_build(&sp->s[0], &ta);
    exc_table_build(&sp->s[1], &tb);
    #pragma omp parallel
    {
        long long ecap = 0;
        ent_t *e = NULL;
        #pragma omp for schedule(dynamic, 64)
        for (long long I = 0; I < N; I++) {
            long long ia, ib;
            det_strings(sp, I, &ia, &ib);
            const exc_t *la = ta.ent + ta.off[ia], *lb = tb.ent + tb.off[ib];
            long long nla = ta.off[ia + 1] - ta.off[ia], nlb = tb.off[ib + 1] - tb.off[ib];
            long long s1a = 0, s1b = 0;
            for (long long x = 0; x < nla; x++) s1a += la[x].deg == 1;
            for (long long y = 0; y < nlb; y++) s1b += lb[y].deg == 1;
            long long bound = nla + nlb + s1a * s1b + 1;
            if (bound > ecap) { free(e); ecap = bound; e = (ent_t *)malloc(sizeof(ent_t) * ecap); }
            long long n = row_entries(sp, ia, ib, la, nla, lb, nlb, m1, m2, e);
            qsort(e, (size_t)n, sizeof(ent_t), cmp_ent);
            double hii = diag_elem(sp->s[0].mask[ia], sp->s[1].mask[ib], sp->norbs, m1, m2);
            for (int k = 0; k < nvec; k++) {
                const double *vk = v + (size_t)k * N;
                double acc = hii * vk[I];
                for (long long t = 0; t < n; t++) acc += e[t].val * vk[e[t].col];
                c[I + (size_t)k * N] = acc;
            }
        }
        free(e);
    }
    exc_table_free(&ta);
    exc_table_free(&tb);
}

/* number of connected off-diagonal pairs in the upper triangle (j > i) */
long long orc_count_upper(const orc_space *sp, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    long long N = sp->ndets, tot = 0;
    exc_table ta, tb;
    exc_table_build(&sp->s[0], &ta);
    exc_table_build(&sp->s[1], &tb);
    #pragma omp parallel for schedule(dynamic, 64) reduction(+:tot)
    for (long long I = 0; I < N; I++) {
        long long ia, ib;
        det_strings(sp, I, &ia, &ib);
        const exc_t *la = ta.ent + ta.off[ia], *lb = tb.ent + tb.off[ib];
        long long nla = ta.off[ia + 1] - ta.off[ia], nlb = tb.off[ib + 1] - tb.off[ib];
        for (long long y = 0; y < nlb; y++) { long long J = det_index(sp, ia, lb[y].idx); tot += J > I; }
        for (long long x = 0; x < nla; x++) {
            long long J = det_index(sp, la[x].idx, ib);
            tot += J > I;
            if (la[x].deg != 1) continue;
            for (long long y = 0; y < nlb; y++) {
                if (lb[y].deg != 1) continue;
                J = det_index(sp, la[x].idx, lb[y].idx);
                tot += J > I;
            }
        }
    }
    exc_table_free(&ta);
    exc_table_free(&tb);
    return tot;
}

/* ---- dense symmetric eigensolver (cyclic Jacobi) --------------------------------
 * Stands in for LAPACK dsyevr (source/mathutil.c:57-109), which is not part of the
 * reference tree.  Eigenvalues ascending; eigenvector i returned in evec[i*n + 0..n-1]. */
int orc_jacobi(int n, const double *a_in, double *eval, double *evec)
{
    double *a = (double *)malloc(sizeof(double) * n * n);
    double *v = (double *)malloc(sizeof(double) * n * n);
    memcpy(a, a_in, sizeof(double) * n * n);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) v[i * n + j] = (i == j);
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0.0, tot = 0.0;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                tot += a[i * n + j] * a[i * n + j];
                if (i != j) off += a[i * n + j] * a[i * n + j];
            }
        if (off <= 1e-30 * (tot > 0 ? tot : 1.0)) break;
        for (int p = 0; p < n - 1; p++)
            for (int q = p + 1; q < n; q++) {
                double apq = a[p * n + q];
                if (fabs(apq) < 1e-300) continue;
                double app = a[p * n + p], aqq = a[q * n + q];
                double theta = (aqq - app) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; k++) {
                    double akp = a[k * n + p], akq = a[k * n + q];
                    a[k * n + p] = c * akp - s * akq;
                    a[k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; k++) {
                    double apk = a[p * n + k], aqk = a[q * n + k];
                    a[p * n + k] = c * apk - s * aqk;
                    a[q * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; k++) {
                    double vkp = v[k * n + p], vkq = v[k * n + q];
                    v[k * n + p] = c * vkp - s * vkq;
                    v[k * n + q] = s * vkp + c * vkq;
                }
            }
    }
    /* sort ascending; columns of v are eigenvectors */
    int *ord = (int *)malloc(sizeof(int) * n);
    for (int i = 0; i < n; i++) ord[i] = i;
    for (int i = 1; i < n; i++) {
        int x = ord[i], j = i - 1;
        while (j >= 0 && a[ord[j] * n + ord[j]] > a[x * n + x]) { ord[j + 1] = ord[j]; j--; }
        ord[j + 1] = x;
    }
    for (int i = 0; i < n; i++) {
        eval[i] = a[ord[i] * n + ord[i]];
        for (int k = 0; k < n; k++) evec[i * n + k] = v[k * n + ord[i]];
    }
    free(ord); free(a); free(v);
    return 0;
}

/* ---- Davidson: source/davidson.c:91-288 control flow ------------------------------
 * One root at a time; residual of root `croot`; new vector -r/(H_II - lambda); MGS;
 * collapse to the krymin lowest Ritz vectors when the space is full or a root
 * converges, then sigma is recomputed for the collapsed space.  Initial guess: the
 * krymin lowest eigenvectors of the leading refdim x refdim block
 * (source/initguess_sbd.c:24-109), or vectors supplied by the caller (guess != NULL,
 * krymin of them).  Returns 0, or -1 if krymin > min(refdim, N) without a guess.
 * cival gets eigenvalue + totfrze; civec[r*N + i] the coefficients; iters_out the
 * number of sigma evaluations of single vectors. */
int orc_davidson(const orc_space *sp, const double *m1, const double *m2, double totfrze,
                 int maxiter, int krymin, int krymax, int nroots, double restol, int refdim,
                 const double *guess, double *cival, double *civec, double *rnorm_out,
                 long long *nsigma_out, int nthreads)
{
    long long N = sp->ndets;
    double *V = (double *)calloc((size_t)krymax * N, sizeof(double));
    double *C = (double *)calloc((size_t)krymax * N, sizeof(double));
    double *hd = (double *)malloc(sizeof(double) * N);
    double *r = (double *)malloc(sizeof(double) * N);
    double *G = (double *)malloc(sizeof(double) * krymax * krymax);
    double *ev = (double *)malloc(sizeof(double) * krymax);
    double *Y = (double *)malloc(sizeof(double) * krymax * krymax);
    long long nsig = 0;
    orc_hdiag(sp, m1, m2, hd, nthreads);
    if (guess) {
        memcpy(V, guess, sizeof(double) * (size_t)krymin * N);
    } else {
        int sb = (int)(refdim < N ? refdim : N);
        if (krymin > sb) { free(V); free(C); free(hd); free(r); free(G); free(ev); free(Y); return -1; }
        double *hb = (double *)malloc(sizeof(double) * sb * sb);
        double *e = (double *)malloc(sizeof(double) * sb), *y = (double *)malloc(sizeof(double) * sb * sb);
        for (int i = 0; i < sb; i++)
            for (int j = i; j < sb; j++) {
                uint64_t ai, bi, aj, bj;
                int d, ph;
                orc_space_det(sp, i, &ai, &bi);
                orc_space_det(sp, j, &aj, &bj);
                hb[i * sb + j] = hb[j * sb + i] = orc_hij(ai, bi, aj, bj, sp->norbs, m1, m2, &d, &ph);
            }
        orc_jacobi(sb, hb, e, y);
        for (int i = 0; i < krymin; i++)
            for (int j = 0; j < sb; j++) V[(size_t)i * N + j] = y[i * sb + j];
        free(hb); free(e); free(y);
    }
    int citer = 1, croot = 1, ckdim = krymin, cflag = 0;
    double rnorm = 0.0;
    while (citer < maxiter && croot <= nroots) {
        orc_sigma(sp, m1, m2, ckdim, V, C, nthreads);
        nsig += ckdim;
        for (int i = 0; i < ckdim; i++)
            for (int j = 0; j < ckdim; j++) {
                double s = 0.0;
                for (long long k = 0; k < N; k++) s += V[(size_t)i * N + k] * C[(size_t)j * N + k];
                G[i * ckdim + j] = s;
            }
        orc_jacobi(ckdim, G, ev, Y);
        while (ckdim < krymax) {
            int rt = croot - 1;
            double lam = ev[rt];
            for (long long k = 0; k < N; k++) {
                double s = 0.0;
                for (int j = 0; j < ckdim; j++)
                    s += Y[rt * ckdim + j] * (C[(size_t)j * N + k] - lam * V[(size_t)j * N + k]);
                r[k] = s;
            }
            double nn = 0.0;
            for (long long k = 0; k < N; k++) nn += r[k] * r[k];
            rnorm = sqrt(nn);
            /* test_convergence: source/davidson.c:408-427 */
            cflag = 0;
            if (rnorm < restol) cflag = croot < nroots ? 1 : 2;
            if (cflag == 2) break;
            if (cflag == 1) { croot++; break; }
            /* generate_newvector: source/davidson.c:293-301 */
            double *nv = V + (size_t)ckdim * N;
            for (long long k = 0; k < N; k++) nv[k] = -r[k] / (hd[k] - lam);
            /* orthonormalize_vector: source/mathutil.c:169-204 (single-pass MGS) */
            for (int i = 0; i < ckdim; i++) {
                double ov = 0.0;
                for (long long k = 0; k < N; k++) ov += V[(size_t)i * N + k] * nv[k];
                for (long long k = 0; k < N; k++) nv[k] -= ov * V[(size_t)i * N + k];
            }
            double nrm = 0.0;
            for (long long k = 0; k < N; k++) nrm += nv[k] * nv[k];
            nrm = sqrt(nrm);
            for (long long k = 0; k < N; k++) nv[k] /= nrm;
            ckdim++;
            orc_sigma(sp, m1, m2, 1, nv, C + (size_t)(ckdim - 1) * N, nthreads);
            nsig += 1;
            for (int i = 0; i < ckdim; i++)
                for (int j = 0; j < ckdim; j++) {
                    double s = 0.0;
                    for (long long k = 0; k < N; k++) s += V[(size_t)i * N + k] * C[(size_t)j * N + k];
                    G[i * ckdim + j] = s;
                }
            orc_jacobi(ckdim, G, ev, Y);
            citer++;
        }
        /* truncate_kspace: source/davidson.c:431-447 */
        double *T = (double *)calloc((size_t)krymin * N, sizeof(double));
        for (int i = 0; i < krymin; i++)
            for (int j = 0; j < ckdim; j++) {
                double y = Y[i * ckdim + j];
                for (long long k = 0; k < N; k++) T[(size_t)i * N + k] += y * V[(size_t)j * N + k];
            }
        memset(V, 0, sizeof(double) * (size_t)krymax * N);
        memcpy(V, T, sizeof(double) * (size_t)krymin * N);
        free(T);
        if (cflag == 2) break;
        ckdim = krymin;
        memset(C, 0, sizeof(double) * (size_t)krymax * N);
    }
    for (int i = 0; i < nroots; i++) cival[i] = ev[i] + totfrze;
    if (civec) memcpy(civec, V, sizeof(double) * (size_t)nroots * N);
    if (rnorm_out) *rnorm_out = rnorm;
    if (nsigma_out) *nsigma_out = nsig;
    free(V); free(C); free(hd); free(r); free(G); free(ev); free(Y);
    return cflag == 2 ? 0 : 1;
}
