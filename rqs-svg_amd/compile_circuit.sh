#!/usr/bin/env bash
# Builds a circuit plug-in for bin/qcs (same contract as the reference compile_circuit.sh:
# a C11 shared object exporting circuit_init / circuit_run, compiled against qcs.h).
# usage: compile_circuit.sh my_circuit.c -o my_circuit.so
set -euo pipefail
: "${CC:=gcc}"
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
"$CC" -fPIC -shared -I"${HERE}/../include" -std=c11 "$@"
