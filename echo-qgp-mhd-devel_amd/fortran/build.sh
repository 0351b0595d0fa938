#!/bin/bash
# Builds the Fortran driver (fortran/echo_amd_driver) against libechoamd.so
# with flang; skips with a message when flang is not installed.
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
FC=${FC:-flang}
if ! command -v "$FC" >/dev/null 2>&1; then FC=/opt/rocm/lib/llvm/bin/flang; fi
if [ ! -x "$FC" ] && ! command -v "$FC" >/dev/null 2>&1; then
  echo "fortran/build.sh: no flang, driver not built"; exit 0
fi
B="$HERE/_build"
mkdir -p "$B"
"$FC" -O2 -J"$B" -c -o "$B/echo_amd_f.o" "$HERE/echo_amd_f.f90"
"$FC" -O2 -J"$B" -c -o "$B/driver.o" "$HERE/driver.f90"
"$FC" -o "$HERE/echo_amd_driver" "$B/echo_amd_f.o" "$B/driver.o" \
  -L"$HERE/.." -lechoamd -Wl,-rpath,'$ORIGIN/..'
echo "build: $HERE/echo_amd_driver"
