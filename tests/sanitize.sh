#!/bin/bash
# The CPU test suite against the host C (mesh, assembly, COO->CSR, drivers) and
# the oracle built with AddressSanitizer and UndefinedBehaviorSanitizer (gcc).
# Device code is not sanitized; GPU tests are not run.  Leak detection is off:
# the interpreter itself keeps allocations until exit.
#   bash tests/sanitize.sh [pytest arguments]
set -e -o pipefail
ROOT=$(cd "$(dirname "$0")/.." && pwd)
make -C "$ROOT/seisfem_amd/csrc" -j8 san
make -C "$ROOT/oracle" san
export SEISFEM_AMD_LIB="$ROOT/seisfem_amd/libseisfem_amd_san.so"
export ORACLE_LIB="$ROOT/oracle/liboracle_san.so"
export LD_PRELOAD="$(gcc -print-file-name=libasan.so) $(gcc -print-file-name=libubsan.so)"
export ASAN_OPTIONS=detect_leaks=0:abort_on_error=1:halt_on_error=1
export UBSAN_OPTIONS=print_stacktrace=1:halt_on_error=1
cd "$ROOT"
python -m pytest tests -q -m "not gpu" -p no:cacheprovider "$@"
