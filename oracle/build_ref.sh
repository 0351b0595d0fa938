#!/bin/bash
# Builds oracle/_ref/libechoref.so: the reference ECHO-QGP hot-path modules,
# compiled unmodified from their sources under /root/reference, plus the
# test harness oracle/ref_harness.f90.  Test infrastructure only: the product
# never links or loads it.
#
# Module order follows the reference makefile's serial build
# (makefile:40, parallel_nompi -> holib -> eos -> common -> viscous ->
# system -> work -> evolve).  init/glaubermc/out/hypersurface/echo are not
# needed: the harness performs the set-up that echo.f90 and init.f90 do.
#
# Flags: -fdefault-real-8 -fdefault-double-8 as the reference makefile
# (makefile:10); -O2 with FP contraction off so that results do not depend
# on FMA formation (the C oracle is built the same way).
set -euo pipefail
REF=${ECHO_REFERENCE:-/root/reference}
HERE="$(cd "$(dirname "$0")" && pwd)"
OUT="$HERE/_ref"
if [ ! -f "$REF/evolve.f90" ]; then
  echo "build_ref: reference sources not found at $REF; skipping" >&2
  exit 0
fi
FC=${FC:-flang}
if ! command -v "$FC" >/dev/null 2>&1; then
  FC=/opt/rocm/lib/llvm/bin/flang
fi
FFLAGS="-fdefault-real-8 -fdefault-double-8 -O2 -fPIC -ffp-contract=off"
mkdir -p "$OUT"
cd "$OUT"
for m in parallel_nompi holib eos common viscous system work evolve hypersurface out; do
  "$FC" $FFLAGS -I"$REF" -c "$REF/$m.f90" -o "$m.o" 2> "$m.log" || { cat "$m.log"; exit 1; }
done
"$FC" $FFLAGS -c "$HERE/ref_harness.f90" -o ref_harness.o 2> ref_harness.log || { cat ref_harness.log; exit 1; }
"$FC" -shared -o libechoref.so parallel_nompi.o holib.o eos.o common.o viscous.o system.o work.o evolve.o \
  hypersurface.o out.o ref_harness.o
echo "build_ref: built $OUT/libechoref.so"
