// SIFS molecular-integral reader: the reference's readmoints (source/readmoints.f90:21-257)
// and the parts of the Columbus colib library it calls, restated in C++.
//
// File layout (Fortran sequential unformatted, 4-byte record markers, 8-byte integers as
// in the reference's GNU build: Makefile -fdefault-integer-8 -frecord-marker=4):
//   record 1   version, ntitle, nsym, nbft, ninfo, nenrgy, nmap           (7 x int64)
//   record 2   title(ntitle)*80, nbpsy(nsym), slabel(nsym)*4, info(ninfo), bfnlab(nbft)*8,
//              ietype(nenrgy), energy(nenrgy) [, imtype(nmap), map(nbft, nmap)]
//              (source/colib/colib8.f:4071-4081, sifzwh)
//   1-e records  l1rec = info(2) words: dword // values(num) // fcore // packed labels
//   2-e records  l2rec = info(4) words: dword // values(num) // packed labels
// with info = (fsplit, l1rec, n1max, l2rec, n2max, ifmt).  The dword packs four 16-bit
// fields (num, lab1, (ibvtyp*8 + itypea)*1024 + itypeb, ifmt*4 + last), most significant
// first (source/colib/colib8.f:337-344 sifd1, :519-526 sifd2).  Labels are an MSB-first
// bit stream in 64-bit words: 8 bits each for ifmt 0 (ulab8, source/colib/colib3.f:1233),
// 10 bits each for ifmt 1 (ulab10, source/colib/colib8.f:4185).
#include "jci_internal.h"

#include <algorithm>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <string>
#include <vector>

namespace {

// one record of a Fortran sequential unformatted file; a record longer than 2^31-1
// bytes comes as subrecords whose leading marker is negative while more follow
bool read_record(FILE *f, std::vector<unsigned char> &out)
{
    out.clear();
    for (;;) {
        int32_t lead, trail;
        if (fread(&lead, 4, 1, f) != 1) return false;
        const size_t n = lead < 0 ? (size_t)(-(int64_t)lead) : (size_t)lead;
        const size_t old = out.size();
        out.resize(old + n);
        if (n && fread(out.data() + old, 1, n, f) != n) return false;
        if (fread(&trail, 4, 1, f) != 1) return false;
        if ((size_t)(trail < 0 ? -(int64_t)trail : trail) != n) return false;
        if (lead >= 0) return true;
    }
}

int64_t get_i8(const unsigned char *p)
{
    int64_t v;
    memcpy(&v, p, 8);
    return v;
}
uint64_t get_u8(const unsigned char *p)
{
    uint64_t v;
    memcpy(&v, p, 8);
    return v;
}

// label k of width w bits in the MSB-first stream starting at word p
long long unpack_bits(const unsigned char *p, long long k, int w)
{
    const long long b = k * w, wi = b / 64;
    const int off = (int)(b % 64);
    const uint64_t hi = get_u8(p + 8 * wi);
    uint64_t v;
    if (off + w <= 64) {
        v = (hi << off) >> (64 - w);
    } else {
        const uint64_t lo = get_u8(p + 8 * (wi + 1));
        const int rest = off + w - 64;
        v = (((hi << off) >> (64 - w)) | (lo >> (64 - rest)));
    }
    return (long long)v;
}

struct Dword {
    long long num, lab1, last, ifmt, itypea, itypeb, ibvtyp;
};

Dword decode_dword(const unsigned char *rec)
{
    Dword d;
    long long u[4];
    for (int i = 0; i < 4; i++) u[i] = unpack_bits(rec, i, 16);
    d.num = u[0];
    d.lab1 = u[1];
    d.last = u[3] % 4;
    d.ifmt = (u[3] / 4) % 8;
    d.itypeb = u[2] % 1024;
    d.itypea = (u[2] / 1024) % 8;
    d.ibvtyp = (u[2] / 8192) % 8;
    return d;
}

// 1-based packed lower-triangle index (readmoints.f90 Ind2Val / Index2E)
long long ind2(long long i, long long j) { return i >= j ? (i - 1) * i / 2 + j : (j - 1) * j / 2 + i; }

struct Sifs {
    FILE *f = nullptr;
    std::string path;
    std::vector<unsigned char> rec;
    long long version = 0, ntitle = 0, nsym = 0, nbft = 0, ninfo = 0, nenrgy = 0, nmap = 0;
    std::vector<long long> nbpsy, info, ietype;
    std::vector<double> energy;
    ~Sifs() { if (f) fclose(f); }

    int fail(const char *msg)
    {
        jci_set_error("readmoints: %s (%s)", msg, path.c_str());
        return -1;
    }

    // headers 1 and 2 (readmoints.f90:96-137)
    int headers()
    {
        if (!read_record(f, rec) || rec.size() < 7 * 8) return fail("*** Error reading header 1! ***");
        long long h[7];
        for (int i = 0; i < 7; i++) h[i] = get_i8(rec.data() + 8 * i);
        version = h[0]; ntitle = h[1]; nsym = h[2]; nbft = h[3]; ninfo = h[4]; nenrgy = h[5]; nmap = h[6];
        if (version != 2) fprintf(stdout, " WARNING: MO's generated with version != 2!\n");
        if (ntitle < 0 || nsym < 1 || nsym > 8 || nbft < 1 || ninfo < 6 || nenrgy < 1 || nmap < 0 ||
            nbft > 1023)
            return fail("*** Error reading header 1! ***");
        if (!read_record(f, rec)) return fail("*** Error reading header 2! ***");
        const size_t need = 80 * ntitle + 8 * nsym + 4 * nsym + 8 * ninfo + 8 * nbft + 16 * nenrgy +
                            (nmap ? 8 * nmap + 8 * nbft * nmap : 0);
        if (rec.size() < need) return fail("*** Error reading header 2! ***");
        size_t o = 80 * ntitle;
        nbpsy.resize(nsym);
        for (long long i = 0; i < nsym; i++, o += 8) nbpsy[i] = get_i8(rec.data() + o);
        o += 4 * nsym;
        info.resize(ninfo);
        for (long long i = 0; i < ninfo; i++, o += 8) info[i] = get_i8(rec.data() + o);
        o += 8 * nbft;
        ietype.resize(nenrgy);
        for (long long i = 0; i < nenrgy; i++, o += 8) ietype[i] = get_i8(rec.data() + o);
        energy.resize(nenrgy);
        for (long long i = 0; i < nenrgy; i++, o += 8) memcpy(&energy[i], rec.data() + o, 8);
        long long tot = 0;
        for (long long n : nbpsy) {
            if (n < 0) return fail("*** Error reading header 2! ***");
            tot += n;
        }
        if (tot != nbft || info[1] < 1 || info[2] < 1 || info[3] < 1 || info[4] < 1)
            return fail("*** Error reading header 2! ***");
        return 0;
    }

    // next integral record of `len` words (seqrbf reads exactly that many)
    int record(long long len, const char *what)
    {
        if (!read_record(f, rec) || (long long)rec.size() < 8 * len) return fail(what);
        return 0;
    }

    // labels of a record: n labels from word lab1 (1-based), bounded by the record length
    int labels(const Dword &d, long long n, long long len, std::vector<long long> &out)
    {
        const int w = d.ifmt == 0 ? 8 : 10;
        const long long words = d.ifmt == 0 ? (n + 7) / 8 : (n + 31) / 32 * 5;
        if (d.lab1 < 1 || d.lab1 - 1 + words > len) return fail("packed labels outside the record");
        out.resize(n);
        const unsigned char *p = rec.data() + 8 * (d.lab1 - 1);
        for (long long k = 0; k < n; k++) out[k] = unpack_bits(p, k, w);
        return 0;
    }
};

}  // namespace

// The reference's readmoints(moints1, moints2, type1n, orbitals, m1len, m2len, en):
// type1n 1 reads H1 (overlap excluded) and the frozen-core energy, 3 reads the three dipole
// arrays into moints1[0..3*m1len), any other value leaves moints1 zero; all but 3 then read
// the 2-e integrals.  en[0] = energy(1) (the nuclear repulsion), en[1] = frozen-core energy.
// Returns 0 or -1 with the reference's message (its Fortran STOP) in jci_last_error().
JCI_API int jci_readmoints(const char *path, int type1n, int orbitals, double *moints1, long long m1len,
                           double *moints2, long long m2len, double *en)
{
    (void)orbitals;   // unused by the reference too (sizes come from the file's nbft)
    Sifs s;
    s.path = path ? path : "moints";
    s.f = fopen(s.path.c_str(), "rb");
    if (!s.f) return s.fail("*** Could not open integral file! ***");
    if (s.headers()) return -1;
    en[0] = s.energy[0];
    en[1] = 0.0;
    const long long nbft = s.nbft, n1int = ind2(nbft, nbft);
    const long long fsplit = s.info[0], l1rec = s.info[1], n1max = s.info[2], l2rec = s.info[3],
                    n2max = s.info[4], ifmt_file = s.info[5];
    const long long mult1 = type1n == 3 ? 3 : 1;
    for (long long i = 0; i < mult1 * m1len; i++) moints1[i] = 0.0;
    std::vector<long long> lab;

    // decode the dword of a 1-e record and check it against the record length
    // (source/colib/colib8.f:333-380, sifd1 with nipv = 2)
    auto decode1 = [&](Dword &d) -> int {
        d = decode_dword(s.rec.data());
        if (d.ifmt != ifmt_file) return s.fail("sifd1: ifmt.ne.info(6)");
        if (d.num < 0 || d.num > n1max) return s.fail("*** Error reading 1-e integrals!***");
        const long long lenpl = d.ifmt == 0 ? (d.num + 3) / 4 : (d.num / 16) * 5;
        const long long lenbv = d.ibvtyp ? (n1max + 63) / 64 : 0;
        if (d.ifmt > 1 || 2 + d.num + lenpl + lenbv > l1rec) return s.fail("*** Error reading 1-e integrals!***");
        if (d.last > 2) return s.fail("*** Error reading 1-e integrals!***");
        return s.labels(d, 2 * d.num, l1rec, lab);
    };
    auto value = [&](long long k) {
        double v;
        memcpy(&v, s.rec.data() + 8 * (1 + k), 8);
        return v;
    };

    if (type1n == 1) {
        // sifrsh -> sifr1n (source/colib/colib8.f:2724-2860, :1858-2200): the itypea = 0
        // arrays; itypeb 0 is the overlap, 1..6 (kinetic, potential, ...) sum into H1,
        // symmetry-blocked lower triangle packed by rows; the frozen-core value of the last
        // record of each array adds to that array's core energy
        std::vector<long long> symb(nbft + 1), nskp(s.nsym), nnskp(s.nsym);
        long long ntot = 0, nntot = 0;
        for (long long i = 0; i < s.nsym; i++) {
            nskp[i] = ntot;
            nnskp[i] = nntot;
            for (long long j = ntot + 1; j <= ntot + s.nbpsy[i]; j++) symb[j] = i;
            ntot += s.nbpsy[i];
            nntot += s.nbpsy[i] * (s.nbpsy[i] + 1) / 2;
        }
        if (nntot > n1int) return s.fail("*** Error reading 1-e integrals!***");
        if (n1int > m1len) return s.fail("moints1 shorter than the file's 1-e array");
        std::vector<double> h1(n1int, 0.0);
        std::vector<long long> offd(s.nsym * s.nsym, 0);
        double fcore_h = 0.0;
        bool any = false;
        Dword d;
        for (;;) {
            if (s.record(l1rec, "*** Error reading 1-e integrals!***")) return -1;
            if (decode1(d)) return -1;
            if (d.itypea == 0 && d.itypeb <= 6) {
                any = true;
                const bool h = d.itypeb >= 1;   // btypes(0:6) = 1, 2, 2, 2, 2, 2, 2
                for (long long k = 0; k < d.num; k++) {
                    long long i = lab[2 * k], j = lab[2 * k + 1];
                    if (i < 1 || j < 1 || i > nbft || j > nbft) return s.fail("1-e label out of range");
                    const long long il = std::max(i, j), jl = std::min(i, j);
                    const long long is = symb[il], js = symb[jl];
                    if (is != js) { offd[is * s.nsym + js]++; continue; }
                    const long long a = il - nskp[is], b = jl - nskp[is];
                    const long long ij = nnskp[is] + a * (a - 1) / 2 + b;
                    if (h) h1[ij - 1] += value(k);
                }
                if (d.last != 0 && h) {
                    double fc;
                    memcpy(&fc, s.rec.data() + 8 * (d.num + 1), 8);
                    fcore_h += fc;
                }
            }
            if (d.last == 2) break;
        }
        if (!any) return s.fail("*** Error reading 1-e integrals!***");
        for (long long v : offd)
            if (v) return s.fail("*** Error reading 1-e integrals!***");   // symmetry blocking errors
        for (long long i = 0; i < n1int; i++) moints1[i] = h1[i];
        en[1] = fcore_h;
    } else if (type1n == 3) {
        // dipole arrays (readmoints.f90:166-205): itypea 1, itypeb 0/1/2 = x/y/z, taken
        // from the records with last = 1; the loop ends at last = 2
        if (3 * n1int > 3 * m1len) return s.fail("moints1 shorter than the file's 1-e array");
        Dword d;
        for (;;) {
            if (s.record(l1rec, "***Error reading integral record!***")) return -1;
            if (decode1(d)) return -1;
            if (d.itypea == 1 && d.itypeb <= 2 && d.last == 1) {
                for (long long k = 0; k < d.num; k++) {
                    const long long i = lab[2 * k], j = lab[2 * k + 1];
                    if (i < 1 || j < 1 || i > nbft || j > nbft) return s.fail("1-e label out of range");
                    moints1[ind2(i, j) - 1 + d.itypeb * m1len] += value(k);
                }
            }
            if (d.last == 2) break;
        }
        return 0;
    } else {
        // no 1-e array requested: skip to the end of the 1-e records below
    }

    // 2-e records (sifo2f -> sifsk1, source/colib/colib8.f:1720-1790, :2870-2960; then
    // readmoints.f90:220-248 with sifd2, colib8.f:447-560)
    if (fsplit != 1) return s.fail("2-e records are separate. Not implemented.");
    if (fseek(s.f, 0, SEEK_SET) != 0 || !read_record(s.f, s.rec) || !read_record(s.f, s.rec))
        return s.fail("*** Error positioning moints2 file! ***");
    for (;;) {
        if (!read_record(s.f, s.rec) || s.rec.size() < 8) return s.fail("*** Error positioning moints2 file! ***");
        const Dword d = decode_dword(s.rec.data());
        if (d.ifmt != ifmt_file) return s.fail("sifd1: ifmt.ne.info(6)");
        if (d.last == 2) break;
        if (d.last > 2) return s.fail("*** Error positioning moints2 file! ***");
    }
    for (long long i = 0; i < m2len; i++) moints2[i] = 0.0;
    for (;;) {
        if (s.record(l2rec, "*** Error reading integral record! ***")) return -1;
        const Dword d = decode_dword(s.rec.data());
        if (d.ifmt != ifmt_file) return s.fail("sifd2: ifmt.ne.info(6)");
        if (d.num < 0 || d.num > n2max || d.ifmt > 1) return s.fail("*** Error unpacking 2-e buffer! ***");
        const long long lenpl = d.ifmt == 0 ? (d.num + 1) / 2 : (d.num / 8) * 5;
        const long long lenbv = d.ibvtyp ? (n2max + 63) / 64 : 0;
        if (1 + d.num + lenpl + lenbv > l2rec) return s.fail("*** Error unpacking 2-e buffer! ***");
        if (d.num > 0 && s.labels(d, 4 * d.num, l2rec, lab)) return -1;
        for (long long k = 0; k < d.num; k++) {
            const long long *L = &lab[4 * k];
            for (int q = 0; q < 4; q++)
                if (L[q] < 1 || L[q] > nbft) return s.fail("2-e label out of range");
            const long long idx = ind2(ind2(L[0], L[1]), ind2(L[2], L[3]));
            if (idx > m2len) return s.fail("moints2 shorter than the file's 2-e array");
            moints2[idx - 1] += value(k);
        }
        if (d.last == 2) break;
        if (d.last > 2) return s.fail("*** Error unpacking 2-e buffer! ***");
    }
    return 0;
}

// replaces source/ioutil.c:427 (readmointegrals): reads the integral file (the reference
// always opens "moints" in the working directory and only prints moflname; here a
// non-empty moflname names the file) and, like the reference's Fortran STOP, ends the
// program with the message on failure
JCI_API void readmointegrals(double *moints1, double *moints2, int itype, int orbitals, char *moflname,
                             int m1len, int m2len, double *nuc_rep, double *fcenergy)
{
    fprintf(stdout, "Calling readmoints_\n");
    fprintf(stdout, " Molecular integral file: %s\n", moflname ? moflname : "");
    fprintf(stdout, " Type of integrals: %lld\n", (long long)itype);
    fprintf(stdout, " 1-e integrals: %lld\n", (long long)m1len);
    fprintf(stdout, " 2-e integrals: %lld\n", (long long)m2len);
    fprintf(stdout, " Reading molecular integrals.\n");
    double en[2] = {0.0, 0.0};
    const char *path = (moflname && *moflname) ? moflname : "moints";
    if (jci_readmoints(path, itype, orbitals, moints1, m1len, moints2, m2len, en) != 0) {
        fprintf(stderr, "%s\n", jci_last_error());
        exit(1);
    }
    *nuc_rep = en[0];
    *fcenergy = en[1];
}
