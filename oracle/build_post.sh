#!/usr/bin/env bash
# Builds the reference's post-processing programs from their own sources
# where they lie under /root/reference, for use as a CHECKER only:
#   avg.x   v1.9/codes/avg_v1.7.f90      time averages -> mean/rms/corr.dat
#   fluc.x  v1.9/codes/fluc.f90          fluctuation field -> fluc.dat
#   qinv.x  v1.8_validation/codes/qInvariant.f90   Q-criterion -> qInvariant.dat
# with the module comdata of v1.9/codes/comdata.f90 (its compile-time nP set
# to the grid, which avg and qInvariant size their arrays with).
#
#   oracle/build_post.sh N1 N2 N3
#
# Output: oracle/_ref/post_n<N1>x<N2>x<N3>/{avg,fluc,qinv}.x (gitignored).
set -euo pipefail
N1=${1:?N1}; N2=${2:?N2}; N3=${3:?N3}
REF=/root/reference
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref/post_n${N1}x${N2}x${N3}"
FC=${FC:-/opt/rocm/lib/llvm/bin/flang}
[ -d "$REF" ] || { echo "reference sources not present: $REF" >&2; exit 3; }
mkdir -p "$OUT"
cd "$OUT"
sed -e "s/=\[130,130,160\]/=[${N1},${N2},${N3}]/" "$REF/v1.9/codes/comdata.f90" > comdata.f90
grep -q "\[${N1},${N2},${N3}\]" comdata.f90 || { echo "nP substitution failed" >&2; exit 4; }
$FC -O2 -c comdata.f90 -o comdata.o
$FC -O2 -o avg.x "$REF/v1.9/codes/avg_v1.7.f90" comdata.o
$FC -O2 -o fluc.x "$REF/v1.9/codes/fluc.f90"
$FC -O2 -o qinv.x "$REF/v1.8_validation/codes/qInvariant.f90" comdata.o
echo "$OUT"
