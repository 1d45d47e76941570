#!/bin/bash
# Builds the compiled-reference cross-check oracle/_ref/ref_instanton from the reference's
# own sources where they lie (instantonmod.f90 and its vendored L-BFGS-B: lbfgsb.f,
# blas.f, linpack.f, timer.f) plus this project's mcmod_mass plugin and harness.
# Flags reproduce ifort -r8 (makefile:5).  detJ's LAPACK call (DSBEVD) is an external
# library the image lacks for Fortran linking, so every routine is emitted in its own
# section and the linker drops the ones the harness never calls (detJ among them).
# TEST INFRASTRUCTURE ONLY; outputs go to oracle/_ref/ (not committed).
set -euo pipefail
REF=${REF:-/root/reference}
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
OUT="$HERE/../_ref"
FC=${FC:-/opt/rocm/lib/llvm/bin/flang}
CC=${CCLANG:-/opt/rocm/lib/llvm/bin/clang}
FFLAGS="-O2 -fdefault-real-8 -fdefault-double-8"
if [ ! -f "$REF/instantonmod.f90" ]; then
  echo "build_ref.sh: $REF/instantonmod.f90 not found; skipping the reference build" >&2
  exit 0
fi
mkdir -p "$OUT"
cd "$OUT"
obj() {  # $1 source -> object with one section per function
  local base
  base=$(basename "$1"); base=${base%.*}
  "$FC" $FFLAGS -S -emit-llvm "$1" -o "$base.ll"
  "$CC" -O2 -c -ffunction-sections -fdata-sections "$base.ll" -o "$base.o"
}
obj "$HERE/ref_plugin.f90"
obj "$REF/instantonmod.f90"
for f in lbfgsb blas linpack timer; do obj "$REF/$f.f"; done
obj "$HERE/ref_harness.f90"
"$FC" ref_harness.o instantonmod.o ref_plugin.o lbfgsb.o blas.o linpack.o timer.o \
  -Wl,--gc-sections -o ref_instanton
rm -f ./*.ll
echo "built $OUT/ref_instanton"
