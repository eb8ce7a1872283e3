#!/bin/bash
# Dataset sweep over the CLI apps, in the manner of the reference's run.sh
# (run.sh:9-16: every dataset of a list, one source per dataset, one app).
#
#   run_sweep.sh [--data DIR] [--suffix .w.gr] [--app bfs|bfs_2s|sssp|pr|cc]
#                [--dry-run] [NAME:SRC ...] [-- extra app flags]
#
# Each NAME:SRC runs  <app> --input DIR/NAME<suffix> --src SRC <extra flags>
# (pr and cc take no source).  Without NAME:SRC pairs, the reference's list
# is used: lj orkut uk-2005 twitter-2010 sk-2005 friendster uk-union with
# sources 0 0 0 0 10 101 1.  A missing input file is reported and skipped.
# --dry-run prints the commands instead of running them.  The exit status
# is the number of runs that failed (0 when all succeeded).
set -u
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
BIN="$HERE/../bin"
DATA="$HOME/data"
SUFFIX=".w.gr"
APP="bfs"
DRY=0
PAIRS=()
EXTRA=()
while [ $# -gt 0 ]; do
  case "$1" in
    --data) DATA="$2"; shift 2 ;;
    --suffix) SUFFIX="$2"; shift 2 ;;
    --app) APP="$2"; shift 2 ;;
    --dry-run) DRY=1; shift ;;
    --) shift; EXTRA=("$@"); break ;;
    *:*) PAIRS+=("$1"); shift ;;
    *) echo "run_sweep: unknown argument $1" >&2; exit 64 ;;
  esac
done
case "$APP" in
  bfs|bfs_2s|sssp|pr|cc) ;;
  *) echo "run_sweep: unknown app $APP" >&2; exit 64 ;;
esac
if [ ${#PAIRS[@]} -eq 0 ]; then
  PAIRS=(lj:0 orkut:0 uk-2005:0 twitter-2010:0 sk-2005:10 friendster:101 uk-union:1)
fi
FAILED=0
for pair in "${PAIRS[@]}"; do
  name="${pair%%:*}"
  src="${pair##*:}"
  input="$DATA/$name$SUFFIX"
  cmd=("$BIN/grus_$APP" --input "$input")
  if [ "$APP" != pr ] && [ "$APP" != cc ]; then cmd+=(--src "$src"); fi
  if [ ${#EXTRA[@]} -gt 0 ]; then cmd+=("${EXTRA[@]}"); fi
  if [ $DRY -eq 1 ]; then
    echo "${cmd[*]}"
    continue
  fi
  if [ ! -f "$input" ]; then
    echo "run_sweep: $input not found, skipped" >&2
    continue
  fi
  echo "== $name (src $src)"
  "${cmd[@]}" || FAILED=$((FAILED + 1))
done
exit $FAILED
