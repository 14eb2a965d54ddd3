#!/bin/bash
# retry: re-run a benchmark/tuning command until its output contains
# ***DONE*** (the reference's retry wrapper, retry:20-28), so that a crash or
# a hang in one trial does not end the search: the journal records the trial
# as BUILD_IN_PROGRESS / RUN_IN_PROGRESS and the next attempt skips it.
# Additions: each attempt runs under a time limit (GAT_RETRY_TIMEOUT seconds,
# default 600) and the number of attempts is capped (GAT_RETRY_MAX, default 20).
set -u
out=$(mktemp "${TMPDIR:-/tmp}/gat_retry.XXXXXX")
limit=${GAT_RETRY_TIMEOUT:-600}
maxn=${GAT_RETRY_MAX:-20}
n=0
rc=1
while [ "$n" -lt "$maxn" ]; do
  n=$((n + 1))
  timeout -k 10 "$limit" "$@" 2>&1 | tee "$out"
  if grep -q '\*\*\*DONE\*\*\*' "$out"; then rc=0; break; fi
  echo "[retry] attempt $n ended without ***DONE***, respawning" >&2
done
rm -f "$out"
exit $rc
