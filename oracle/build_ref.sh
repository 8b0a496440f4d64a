#!/usr/bin/env bash
# Builds the reference CPU path from its own headers, where they lie, into
# oracle/_ref/ (never copied into this repository).  Needs the MPICH and
# MKL (LAPACK dsyev) that ship in /opt/conda of this image.
#   oracle/_ref/ref_driver      default CPU sigma (chemistry/tpb/mult.h:625-881)
#   oracle/_ref/ref_driver_trad SBD_TRADMODE sigma (mult.h:113-621)
#   oracle/_ref/diag            the reference CLI app (apps/chemistry_tpb_selected_basis_diagonalization)
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
REF="${SBD_REFERENCE:-/root/reference}"
OUT="$HERE/_ref"
if [ ! -d "$REF/include/sbd" ]; then
  echo "reference tree not found at $REF; skipping oracle/_ref" >&2
  exit 0
fi
mkdir -p "$OUT"
export MPICH_CXX=g++
CXX=/opt/conda/bin/mpicxx
FLAGS=(-std=c++17 -O3 -fopenmp -w -I"$REF/include" -I/opt/conda/include)
LIBS=(-static-libstdc++ -static-libgcc /opt/conda/lib/libmkl_rt.so -Wl,-rpath,/opt/conda/lib)
PRE=(-include iostream -include chrono -include fstream -include random -include deque)
$CXX "${FLAGS[@]}" "${PRE[@]}" "$HERE/ref_driver.cc" -o "$OUT/ref_driver" "${LIBS[@]}"
$CXX "${FLAGS[@]}" "${PRE[@]}" -DSBD_TRADMODE "$HERE/ref_driver.cc" -o "$OUT/ref_driver_trad" "${LIBS[@]}"
$CXX "${FLAGS[@]}" -DSBD_TRADMODE \
  "$REF/apps/chemistry_tpb_selected_basis_diagonalization/main.cc" -o "$OUT/diag" "${LIBS[@]}"
echo "built: $(ls "$OUT")"
