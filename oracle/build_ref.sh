#!/usr/bin/env bash
# Compiles the unmodified reference program (Fortran 2003 + OpenMP, one C
# helper) from the sources where they lie under /root/reference/src into
# oracle/_ref/FSILBM3D.  The reference's own Makefile targets gfortran/ifort,
# which this image lacks; amdflang (LLVM flang, ROCm) builds the same sources.
#   -ffp-contract=off : no fused multiply-add contraction, so every product and
#                       sum is rounded the way the source spells it.
# The module order follows the reference Makefile's SRC list (Makefile:33).
# Outputs go only to oracle/_ref/ (git-ignored; it travels to the GPU host as a
# built binary and is run there as the CPU baseline and as the parity oracle).
set -euo pipefail
REF=${FSIL_REFERENCE:-/root/reference}
HERE="$(cd "$(dirname "$0")" && pwd)"
OUT="$HERE/_ref"
OBJ="$OUT/obj"
if [ ! -d "$REF/src" ]; then
  echo "build_ref.sh: $REF/src not found; keeping any prebuilt $OUT/FSILBM3D" >&2
  exit 0
fi
FC=${FC:-amdflang}
FFLAGS="-O3 -fopenmp -ffree-form -ffp-contract=off"
mkdir -p "$OBJ"
for m in ConstParams FlowCondition SolidSolver Solidbody FluidDomain LBMBlockComm Util main; do
  "$FC" $FFLAGS -module-dir "$OBJ" -c "$REF/src/$m.f90" -o "$OBJ/$m.o"
done
gcc -O3 -c "$REF/src/forkthread.c" -o "$OBJ/forkthread.o"
"$FC" -fopenmp -o "$OUT/FSILBM3D.tmp" "$OBJ"/ConstParams.o "$OBJ"/FlowCondition.o "$OBJ"/SolidSolver.o \
  "$OBJ"/Solidbody.o "$OBJ"/FluidDomain.o "$OBJ"/LBMBlockComm.o "$OBJ"/Util.o "$OBJ"/main.o "$OBJ"/forkthread.o
mv "$OUT/FSILBM3D.tmp" "$OUT/FSILBM3D"
echo "built $OUT/FSILBM3D"
