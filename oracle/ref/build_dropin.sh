#!/bin/bash
# Links the reference's own RPI driver, rpi_ser.f90 (compiled from its source where it lies,
# with the reference's -r8 flags, makefile:5), against the drop-in Fortran modules of
# pimd-tunneling_amd/fortran and libpimdtunnel.so, once per kind of plugin:
#   oracle/_ref/rpi_dropin_device  the device plugin mcmod_rp.f90 + rp_bridge_device.f90
#   oracle/_ref/rpi_dropin_host    a host plugin written to the reference's plugin API
#                                  (ref_plugin.f90: the config potentials in plain Fortran)
#                                  + rp_bridge_host.f90, which calls it back per chain
# The two programs show the link a maintainer makes and let a GPU test run
# the reference's driver on the device.  rpi_ser.f90's DSYEVD call (under checkhess) is
# resolved from the MKL runtime in /opt/conda/lib.  Needs libpimdtunnel.so built first.
# TEST INFRASTRUCTURE ONLY; outputs go to oracle/_ref/ (not committed).
set -euo pipefail
REF=${REF:-/root/reference}
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
ROOT="$HERE/../.."
PKG="$ROOT/pimd-tunneling_amd"
OUT="$HERE/../_ref"
FC=${FC:-/opt/rocm/lib/llvm/bin/flang}
MKLDIR=${MKLDIR:-/opt/conda/lib}
if [ ! -f "$REF/rpi_ser.f90" ]; then
  echo "build_dropin.sh: $REF/rpi_ser.f90 not found; skipping the drop-in link" >&2
  exit 0
fi
if [ ! -f "$PKG/lib/libpimdtunnel.so" ]; then
  echo "build_dropin.sh: build libpimdtunnel.so first" >&2
  exit 1
fi
mkdir -p "$OUT"
OUT="$(cd "$OUT" && pwd)"
F="$PKG/fortran"
link() {  # $1 name, $2 plugin source, $3 bridge source
  local dir="$OUT/build_$1"
  rm -rf "$dir"
  mkdir -p "$dir"
  cd "$dir"
  "$FC" -O2 -module-dir . -c "$F/rp_hip.f90" -o rp_hip.o
  "$FC" -O2 -module-dir . -c "$2" -o plugin.o
  "$FC" -O2 -module-dir . -c "$3" -o bridge.o
  "$FC" -O2 -module-dir . -c "$F/instantonmod_rp.f90" -o instantonmod.o
  "$FC" -O2 -module-dir . -c "$F/verletint_rp.f90" -o verletint.o
  "$FC" -O2 -fdefault-real-8 -fdefault-double-8 -module-dir . -c "$REF/rpi_ser.f90" -o rpi_ser.o
  "$FC" rpi_ser.o verletint.o instantonmod.o bridge.o plugin.o rp_hip.o \
    -L"$PKG/lib" -lpimdtunnel -L"$MKLDIR" -lmkl_rt \
    -Wl,-rpath,'$ORIGIN/../../pimd-tunneling_amd/lib' -Wl,-rpath,"$MKLDIR" -o "$OUT/$1"
  cd "$OUT"
  rm -rf "$dir"
  echo "built $OUT/$1"
}
link rpi_dropin_device "$F/mcmod_rp.f90" "$F/rp_bridge_device.f90"
link rpi_dropin_host "$HERE/ref_plugin.f90" "$F/rp_bridge_host.f90"
