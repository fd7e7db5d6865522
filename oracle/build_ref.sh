#!/bin/bash
# Builds the reference TAURUS_vap (Fortran) from its own sources, where they lie
# under /root/reference/src, into oracle/_ref/.  Test infrastructure only: the
# product never links or runs anything built here.
#
# Outputs (oracle/_ref/, git-ignored, shipped with the tree):
#   taurus_vap32.exe   stock reference, TBMEs stored as real32 (rH2 = real32)
#   taurus_vap64.exe   same sources with the kind rH2 = real64 (the "FP64 oracle")
#   harness32.exe      reference modules + oracle/ref_harness.f90 (stage dumper)
#   harness64.exe      same, rH2 = real64
#
# Toolchain: AMD flang (ROCm) with OpenMP.  BLAS/LAPACK: the image has no system
# liblapack/libblas; the LP64 OpenBLAS shipped inside scipy exports every routine
# under a "scipy_" prefix, so the undefined BLAS/LAPACK references of the objects
# built here are renamed with objcopy --redefine-syms (names taken from nm of our
# own objects and from the library's exported-symbol table only).
#
# The one source transformation is the kind override for the FP64 build
# (module_constants.f90: "rH2 = real32" -> "rH2 = real64"), applied to a scratch
# copy outside the repository, the same way the reference's own build applies its
# "!cmpi" switch (compile.sh:124-137).  Nothing from /root/reference is written
# into the repository.
set -euo pipefail

REFSRC=${REFSRC:-/root/reference/src}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT=$HERE/_ref
FC=${FC:-flang}
OPT=${OPT:-"-O3 -fopenmp"}
OPENBLAS=${OPENBLAS:-$(python3 -c 'import glob,scipy,os; print(glob.glob(os.path.join(os.path.dirname(scipy.__file__), "..", "scipy.libs", "libscipy_openblas*.so"))[0])')}
OPENBLAS=$(readlink -f "$OPENBLAS")

if [ ! -d "$REFSRC" ]; then
  echo "build_ref.sh: reference sources not found at $REFSRC (nothing to build)" >&2
  exit 3
fi
command -v "$FC" >/dev/null || { echo "build_ref.sh: $FC not found" >&2; exit 3; }

mkdir -p "$OUT"
SCRATCH=$(mktemp -d /tmp/taurus_ref_build.XXXXXX)
trap 'rm -rf "$SCRATCH"' EXIT

MODS="module_constants module_mathmethods module_nucleus module_basis \
      module_hamiltonian module_wavefunctions module_fields module_particlenumber \
      module_pairs module_angularmomentum module_multipoles module_radius \
      module_operators module_projection module_constraints module_gradient \
      module_initialization"

build_variant () {
  local tag=$1 kind=$2
  local d=$SCRATCH/$tag
  mkdir -p "$d"
  for m in $MODS; do cp "$REFSRC/$m.f90" "$d/"; done
  cp "$REFSRC/subroutines_pfaffian.f" "$REFSRC/taurus_vap.f90" "$d/"
  cp "$HERE/ref_harness.f90" "$d/"
  if [ "$kind" = real64 ]; then
    sed -i 's/rH2 = real32/rH2 = real64/' "$d/module_constants.f90"
    grep -q 'rH2 = real64' "$d/module_constants.f90"
  fi
  ( cd "$d"
    for m in $MODS; do $FC $OPT -c $m.f90; done
    $FC $OPT -c subroutines_pfaffian.f
    $FC $OPT -c taurus_vap.f90
    $FC $OPT -c ref_harness.f90
    # rename BLAS/LAPACK references to the scipy-prefixed exports
    nm -D --defined-only "$OPENBLAS" | awk '$3 ~ /^scipy_/ {print $3}' | sed 's/^scipy_//' | sort -u > exported
    nm -u *.o | awk '{print $NF}' | sort -u > undef
    comm -12 undef exported | awk '{print $1" scipy_"$1}' > redefine
    for o in *.o; do objcopy --redefine-syms=redefine "$o"; done
    local objs=""
    for m in $MODS; do objs="$objs $m.o"; done
    $FC $OPT -o "$OUT/taurus_vap$tag.exe" $objs subroutines_pfaffian.o taurus_vap.o \
        "$OPENBLAS" -Wl,-rpath,"$(dirname "$OPENBLAS")"
    $FC $OPT -o "$OUT/harness$tag.exe" $objs subroutines_pfaffian.o ref_harness.o \
        "$OPENBLAS" -Wl,-rpath,"$(dirname "$OPENBLAS")"
  )
}

build_variant 32 real32
build_variant 64 real64
echo "build_ref.sh: built $(ls "$OUT" | tr '\n' ' ')"
