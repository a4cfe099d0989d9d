#!/bin/sh
# Build oracle/_ref/sifs_ref, the SIFS file writer of the tests: colib's header writer,
# label / dword packers and record writer (source/colib/colib{1,2,3,7,8}.f, compiled from
# where they lie under the reference tree) with the harness oracle/sifs/sifs_ref.f90.
# Integer kind follows the reference's GNU build flags (Makefile: -fdefault-integer-8
# -frecord-marker=4; flang's sequential records carry 4-byte markers).  Test
# infrastructure only.
#
# Not built: the reference reader readmoints (source/readmoints.f90) and colib's record
# decoders, which copy values through colib's dcopy_wr -> BLAS dcopy; the image has no
# Fortran BLAS, so that path is unbuildable here.
set -e
REF=${JAYCI_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/../_ref"
FC=${FC:-/opt/rocm/bin/amdflang}
[ -d "$REF/source/colib" ] || { echo "reference tree not found: $REF" >&2; exit 2; }
mkdir -p "$OUT/sifs_obj"
for f in colib1 colib2 colib3 colib7 colib8; do
  $FC -c -O1 -w -fdefault-integer-8 "$REF/source/colib/$f.f" -o "$OUT/sifs_obj/$f.o"
done
$FC -c -O1 -w -fdefault-integer-8 "$HERE/sifs_ref.f90" -o "$OUT/sifs_obj/sifs_ref.o"
# The colib files also hold routines unrelated to SIFS that call BLAS/LAPACK and other
# colib files; those names stay unresolved (nothing is linked in their place).  The
# harness never reaches them; the list is kept for the record.
$FC -Wl,--unresolved-symbols=ignore-all -o "$OUT/sifs_ref" "$OUT/sifs_obj"/*.o
nm -u "$OUT/sifs_ref" | awk '{print $2}' | grep '_$' | tr '\n' ' ' > "$OUT/sifs_unresolved.txt"
echo "unresolved Fortran names (unreachable from the harness): $(cat "$OUT/sifs_unresolved.txt")"
echo "$OUT/sifs_ref"
