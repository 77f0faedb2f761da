#!/bin/sh
# Build the high-precision ERI oracle (test infrastructure) next to this script.
set -e
cd "$(dirname "$0")"
gcc -O2 -fPIC -shared -std=gnu11 -o liboracle_hp.so.tmp eri_hp.c -lquadmath -lm
mv liboracle_hp.so.tmp liboracle_hp.so
