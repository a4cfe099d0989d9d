// Determinant-space descriptors: electron-occupation groups per spin and the allowed
// (p,q) group blocks, in the reference's order.
//
//   groups .... source/citruncate.c:205-251 (docc descending, actv descending, virt ascending)
//   strings ... source/citruncate.c:321-415 (docc sub-string outermost, then actv, then virt,
//               each in lexicographic order, source/straddress.c:77)
//   blocks .... serial rule source/citruncate.c:291-312; MRCI-SD rule
//               mpi_source/citruncate.c:414-471; (p,q) order p-major, alpha-major inside
//               (source/citruncate.c:497-538)
#include "jci_internal.h"
#include <string.h>

static int imin(int a, int b) { return a < b ? a : b; }
static int imax(int a, int b) { return a > b ? a : b; }

static long long nck(int n, int k)
{
    if (k < 0 || n < 0 || k > n) return 0;
    long long r = 1;
    for (int i = 1; i <= k; i++) r = r * (n - k + i) / i;
    return r;
}

void jci_binom_table(unsigned long long *tab)
{
    for (int n = 0; n <= 64; n++)
        for (int k = 0; k <= 64; k++) {
            if (k > n) { tab[n * 65 + k] = 0; continue; }
            if (k == 0 || k == n) { tab[n * 65 + k] = 1; continue; }
            unsigned long long a = tab[(n - 1) * 65 + k - 1], b = tab[(n - 1) * 65 + k];
            tab[n * 65 + k] = (a + b < a) ? ~0ULL : a + b;   // saturate (never used that large)
        }
}

static int spin_desc(int nelec, int norbs, int ndocc, int nactv, int xlvl, SpinDesc *s)
{
    memset(s, 0, sizeof(*s));
    s->nelec = nelec;
    int vorbs = norbs - ndocc - nactv;
    int min_d = ndocc - imin(ndocc, xlvl), max_d = ndocc;
    int min_a = (nelec - ndocc) - imin(nelec - ndocc, xlvl);
    int max_a = imin(imin(ndocc, xlvl) + (nelec - ndocc), nactv);
    int n = 0;
    long long start = 0;
    for (int i = max_d; i >= min_d; i--)
        for (int j = max_a; j >= min_a; j--)
            for (int k = 0; k <= imin(vorbs, xlvl); k++) {
                if (i + j + k != nelec) continue;
                if (n >= JCI_MAXG) {
                    jci_set_error("space: more than %d electron-occupation groups", JCI_MAXG);
                    return -1;
                }
                s->gd[n] = i; s->ga[n] = j; s->gv[n] = k;
                s->nsa[n] = nck(nactv, j);
                s->nsv[n] = nck(vorbs, k);
                s->gn[n] = nck(ndocc, i) * s->nsa[n] * s->nsv[n];
                s->gstart[n] = start;
                start += s->gn[n];
                n++;
            }
    s->gstart[n] = start;
    s->ngrp = n;
    return 0;
}

int jci_build_desc(int na, int nb, int norbs, int ndocc, int nactv, int xlvl, int rule,
                   SpaceDesc *d)
{
    memset(d, 0, sizeof(*d));
    if (norbs < 1 || norbs > 64) { jci_set_error("space: norbs must be in 1..64"); return -1; }
    if (ndocc < 0 || nactv < 0 || ndocc + nactv > norbs) {
        jci_set_error("space: need 0 <= ndocc, nactv and ndocc + nactv <= norbs");
        return -1;
    }
    if (na < 0 || nb < 0 || na > norbs || nb > norbs) {
        jci_set_error("space: electron counts out of range");
        return -1;
    }
    d->norbs = norbs; d->ndocc = ndocc; d->nactv = nactv; d->nvirt = norbs - ndocc - nactv;
    d->dmask = ndocc ? ((1ULL << ndocc) - 1ULL) : 0ULL;
    unsigned long long inner = (ndocc + nactv) >= 64 ? ~0ULL : ((1ULL << (ndocc + nactv)) - 1ULL);
    d->amask = inner & ~d->dmask;
    d->allmask = norbs >= 64 ? ~0ULL : ((1ULL << norbs) - 1ULL);
    d->vmask = d->allmask & ~inner;
    if (spin_desc(na, norbs, ndocc, nactv, xlvl, &d->s[0]) ||
        spin_desc(nb, norbs, ndocc, nactv, xlvl, &d->s[1]))
        return -1;
    int doccmin, actvmin;
    if (rule == 0) { doccmin = 2 * (ndocc - imin(ndocc, xlvl)); actvmin = -1; }
    else {
        doccmin = imax(2 * ndocc - xlvl, 0);
        actvmin = imax((na - ndocc) + (nb - ndocc) - xlvl, 0);
    }
    long long off = 0;
    int nbk = 0;
    for (int p = 0; p < JCI_MAXG; p++)
        for (int q = 0; q < JCI_MAXG; q++) d->blkid[p][q] = -1;
    for (int p = 0; p < d->s[0].ngrp; p++)
        for (int q = 0; q < d->s[1].ngrp; q++) {
            if (d->s[0].gd[p] + d->s[1].gd[q] < doccmin) continue;
            if (rule != 0 && d->s[0].ga[p] + d->s[1].ga[q] < actvmin) continue;
            if (d->s[0].gv[p] + d->s[1].gv[q] > xlvl) continue;
            if (d->s[0].gn[p] == 0 || d->s[1].gn[q] == 0) continue;
            d->bp[nbk] = p; d->bq[nbk] = q; d->boff[nbk] = off;
            d->blkid[p][q] = nbk;
            off += d->s[0].gn[p] * d->s[1].gn[q];
            nbk++;
        }
    d->boff[nbk] = off;
    d->nblk = nbk;
    d->ndets = off;
    return 0;
}

// The space with its spins swapped: spin 0 describes the beta strings, blocks keep their
// order and offsets (block b holds the same determinants, stored beta-major).
void jci_transpose_desc(const SpaceDesc &D, SpaceDesc *T)
{
    *T = D;
    T->s[0] = D.s[1];
    T->s[1] = D.s[0];
    for (int i = 0; i < JCI_MAXG; i++)
        for (int j = 0; j < JCI_MAXG; j++) T->blkid[i][j] = -1;
    for (int b = 0; b < D.nblk; b++) {
        T->bp[b] = D.bq[b];
        T->bq[b] = D.bp[b];
        T->blkid[D.bq[b]][D.bp[b]] = b;
    }
}
