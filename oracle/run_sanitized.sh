#!/bin/bash
# Runs the CPU test suite with the oracle loaded from its AddressSanitizer + UBSan build (host code;
# only oracle.cpp is instrumented, no GPU code is built or run).  Usage: bash oracle/run_sanitized.sh
set -e
cd "$(dirname "$0")/.."
make -s -C oracle sanitize
ASAN_LIB=$(g++ -print-file-name=libasan.so)
UBSAN_LIB=$(g++ -print-file-name=libubsan.so)
LD_PRELOAD="$ASAN_LIB:$UBSAN_LIB" ASAN_OPTIONS=detect_leaks=0:abort_on_error=1 UBSAN_OPTIONS=print_stacktrace=1:halt_on_error=1 \
  BH_ORACLE_SANITIZE=1 python -m pytest -q -m "not gpu" -p no:cacheprovider tests/ "$@"
