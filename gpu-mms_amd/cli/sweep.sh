#!/bin/bash
# Size sweeps and repeat runs of testMM, the counterparts of the reference's
# scripts/allSizes.sh, scripts/powersOfTwoInput.sh and scripts/raceConditions.sh.
#   sweep.sh sizes  [reps]   N = 2^k + a random fraction of 2^k, k = 10..26
#   sweep.sh pow2            N = 2^k, k = 10..28
#   sweep.sh repeat [count]  N = 2^20 + 9000, count runs (default 1000)
# Each run prints N, then testMM's "avg min max" line and its verdicts.
set -u
HERE="$(cd "$(dirname "$0")" && pwd)"
T="$HERE/testMM"
mode="${1:-sizes}"
run() { echo "N=$1"; "$T" "$1" || { echo "testMM failed for N=$1"; exit 1; }; }
case "$mode" in
  sizes)
    reps="${2:-10}"
    for k in $(seq 10 26); do
      for _ in $(seq 1 "$reps"); do run $(( (1 << k) + RANDOM * (1 << k) / 32768 )); done
    done ;;
  pow2)
    for k in $(seq 10 28); do run $(( 1 << k )); done ;;
  repeat)
    count="${2:-1000}"
    for _ in $(seq 1 "$count"); do run $(( (1 << 20) + 9000 )); done ;;
  *) echo "usage: $0 sizes [reps] | pow2 | repeat [count]"; exit 2 ;;
esac
