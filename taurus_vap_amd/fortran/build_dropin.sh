#!/bin/bash
# Integration build: the reference TAURUS_vap driver with the GPU projection.
#
# Compiles the reference sources (read from REFSRC, default /root/reference/src)
# together with taurus_gpu.f90 and links libtaurus_gpu.so.  The reference's
# module_projection.f90 gets the two-line change a maintainer makes to adopt the
# library: its project_wavefunction is removed and project_wavefunction_gpu.inc is
# included in the module in its place.  The edited copy lives in a scratch directory; outputs go to
# taurus_vap_amd/fortran/_build/ (taurus_vap_gpu32.exe: stock rH2 = real32,
# taurus_vap_gpu64.exe: rH2 = real64).
set -euo pipefail

REFSRC=${REFSRC:-/root/reference/src}
HERE=$(cd "$(dirname "$0")" && pwd)
PKG=$(cd "$HERE/.." && pwd)
OUT=$HERE/_build
FC=${FC:-flang}
OPT=${OPT:-"-O3 -fopenmp"}
OPENBLAS=${OPENBLAS:-$(python3 -c 'import glob,scipy,os; print(glob.glob(os.path.join(os.path.dirname(scipy.__file__), "..", "scipy.libs", "libscipy_openblas*.so"))[0])')}
OPENBLAS=$(readlink -f "$OPENBLAS")

[ -d "$REFSRC" ] || { echo "build_dropin.sh: reference sources not found at $REFSRC" >&2; exit 3; }
[ -f "$PKG/libtaurus_gpu.so" ] || { echo "build_dropin.sh: build libtaurus_gpu.so first" >&2; exit 3; }
mkdir -p "$OUT"
SCRATCH=$(mktemp -d /tmp/taurus_dropin.XXXXXX)
trap 'rm -rf "$SCRATCH"' EXIT

PRE="module_constants module_mathmethods module_nucleus module_basis module_hamiltonian \
     module_wavefunctions module_fields"
POST="module_particlenumber module_pairs module_angularmomentum module_multipoles \
      module_radius module_operators module_projection module_constraints module_gradient \
      module_initialization"

build_variant () {
  local tag=$1 kind=$2 d=$SCRATCH/$1
  mkdir -p "$d"
  for m in $PRE $POST; do cp "$REFSRC/$m.f90" "$d/"; done
  cp "$REFSRC/subroutines_pfaffian.f" "$REFSRC/taurus_vap.f90" "$d/"
  cp "$HERE/taurus_gpu.f90" "$HERE/project_wavefunction_gpu.inc" "$d/"
  if [ "$kind" = real64 ]; then
    sed -i 's/rH2 = real32/rH2 = real64/' "$d/module_constants.f90"
  fi
  # adopt the library in MODULE Projection: the reference's project_wavefunction
  # is removed and the drop-in included in its place
  sed -i -e '/^subroutine project_wavefunction(iopt,iprint,Mphip,Mphin,ndim)/,/^end subroutine project_wavefunction *$/d' \
         -e "s/^END MODULE Projection/include 'project_wavefunction_gpu.inc'\n\nEND MODULE Projection/" \
         "$d/module_projection.f90"
  [ "$(grep -c '^subroutine project_wavefunction(' "$d/module_projection.f90")" = 0 ] || { echo "patch failed" >&2; exit 4; }
  grep -q "include 'project_wavefunction_gpu.inc'" "$d/module_projection.f90" || { echo "patch failed" >&2; exit 4; }
  # and in MODULE Gradient: diagonalize_hsp_and_H11(0) on the device, the
  # reference's routine kept under another name for opt = 1 (eigenbasis files)
  cp "$HERE/diagonalize_hsp_and_H11_gpu.inc" "$d/"
  sed -i -e 's/^subroutine diagonalize_hsp_and_H11(opt,ndim)/subroutine diagonalize_hsp_and_H11_files(opt,ndim)/' \
         -e 's/^end subroutine diagonalize_hsp_and_H11 *$/end subroutine diagonalize_hsp_and_H11_files/' \
         -e "s/^END MODULE Gradient/include 'diagonalize_hsp_and_H11_gpu.inc'\n\nEND MODULE Gradient/" \
         "$d/module_gradient.f90"
  [ "$(grep -c 'diagonalize_hsp_and_H11_files' "$d/module_gradient.f90")" = 2 ] || { echo "patch failed" >&2; exit 4; }
  grep -q "include 'diagonalize_hsp_and_H11_gpu.inc'" "$d/module_gradient.f90" || { echo "patch failed" >&2; exit 4; }
  ( cd "$d"
    for m in $PRE; do $FC $OPT -c $m.f90; done
    $FC $OPT -c taurus_gpu.f90
    for m in $POST; do $FC $OPT -c $m.f90; done
    $FC $OPT -c subroutines_pfaffian.f taurus_vap.f90
    nm -D --defined-only "$OPENBLAS" | awk '$3 ~ /^scipy_/ {print $3}' | sed 's/^scipy_//' | sort -u > exported
    nm -u *.o | awk '{print $NF}' | sort -u > undef
    comm -12 undef exported | awk '{print $1" scipy_"$1}' > redefine
    for o in *.o; do objcopy --redefine-syms=redefine "$o"; done
    objs=""
    for m in $PRE; do objs="$objs $m.o"; done
    objs="$objs taurus_gpu.o"
    for m in $POST; do objs="$objs $m.o"; done
    $FC $OPT -o "$OUT/taurus_vap_gpu$tag.exe" $objs subroutines_pfaffian.o taurus_vap.o \
        "$PKG/libtaurus_gpu.so" "$OPENBLAS" \
        -Wl,-rpath,'$ORIGIN/../..' -Wl,-rpath,"$(dirname "$OPENBLAS")"
  )
}

build_variant 32 real32
build_variant 64 real64
echo "build_dropin.sh: built $(ls "$OUT" | tr '\n' ' ')"
