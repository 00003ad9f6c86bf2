#!/usr/bin/env bash
# Builds the reference solver (halo-corrected v1.9 physics) from its own
# sources where they lie under /root/reference, for use as a CHECKER only.
#
#   oracle/build_ref.sh N1 N2 N3 MAXSTEP NWRITE [STUDY]
#
# Output: oracle/_ref/n<N1>x<N2>x<N3>_s<MAXSTEP>_w<NWRITE>[_STUDY]/main.x
# (gitignored).
#
# STUDY names one of the reference's wall-condition studies: s1..s6
# (v1.9_STWNV), rd1..rd3 (v1.9_RD), g1 (v1.9_groov), gst1..gst5
# (v1.9_groov_strat), st1..st4 (v1.9_strat).  They differ from v1.9 only in
# the bottom-wall block of the boundary update.  rd1..rd3 already carry the
# halo-corrected arrays and are built from their own main and comdata.  The
# others keep v1.9's halo-less var array, whose 4th-order stencil reads out
# of bounds (main_v1.9_s1.f90:38); for them the study's bottom-wall block is
# spliced into the halo-corrected main, so that the build differs from the
# canonical physics only in the wall condition under study.
#
# The reference sets its grid and step count as compile-time parameters
# (v1.9/codes/comdata.f90:4,8-9), so the build substitutes them in a copy
# written under oracle/_ref (never committed).  Two further build-time edits
# instrument the output only: the 3D solution file is written every NWRITE
# steps instead of every 100 (main_v1.9.f90:769) and with ES25.17 instead of
# E22.15 (main_v1.9.f90:788) so that a file holds every bit of the state.
# The numerics are untouched.  With REF_FMT=E22.15 in the environment the
# files keep the reference's own E22.15 format (output directory suffix
# _e22), so that the build can restart (resOp=2) from the files it wrote,
# whose reader expects that format.  MPI is the image's MPICH 3.3.2; its Fortran
# module was built by gfortran and cannot be read by flang, so `use mpi` is
# satisfied by compiling MPICH's own mpif.h into a module (what MPICH's
# mpi module is made from); at one rank every exchange is a self-copy.
set -euo pipefail
N1=${1:?N1}; N2=${2:?N2}; N3=${3:?N3}; MAXSTEP=${4:?MAXSTEP}; NWRITE=${5:-100}; STUDY=${6:-}
REF=/root/reference
HERE=$(cd "$(dirname "$0")" && pwd)
FMTSUF=""
[ "${REF_FMT:-}" = "E22.15" ] && FMTSUF="_e22"
OUT="$HERE/_ref/n${N1}x${N2}x${N3}_s${MAXSTEP}_w${NWRITE}${STUDY:+_$STUDY}${FMTSUF}"
MAINSRC="$REF/v1.10/codes/main_v1.9.f90"
COMSRC="$REF/v1.9/codes/comdata.f90"
SPLICE=""
case "$STUDY" in
  "") ;;
  rd[1-3]) MAINSRC="$REF/v1.9_RD/v1.9_$STUDY/codes/main_v1.9_$STUDY.f90"
           COMSRC="$REF/v1.9_RD/v1.9_$STUDY/codes/comdata.f90" ;;
  s[1-6]) SPLICE="$REF/v1.9_STWNV/v1.9_$STUDY/codes/main_v1.9_$STUDY.f90" ;;
  g1) SPLICE="$REF/v1.9_groov/v1.9_g1/codes/main_v1.9_g1.f90" ;;
  gst[1-5]) SPLICE="$REF/v1.9_groov_strat/v1.9_$STUDY/codes/main_v1.9_$STUDY.f90" ;;
  st[1-4]) SPLICE="$REF/v1.9_strat/v1.9_$STUDY/codes/main_v1.9_$STUDY.f90" ;;
  *) echo "unknown study: $STUDY" >&2; exit 2 ;;
esac
FC=${FC:-/opt/rocm/lib/llvm/bin/flang}
MPIINC=/opt/conda/include
MPILIB=/opt/conda/lib
[ -d "$REF" ] || { echo "reference sources not present: $REF" >&2; exit 3; }
mkdir -p "$OUT"
cd "$OUT"

sed -e "s/=\[130,130,160\]/=[${N1},${N2},${N3}]/" \
    -e "s/maxstep\t=int(5E7)/maxstep\t=${MAXSTEP}/" \
    "$COMSRC" > comdata.f90
grep -q "\[${N1},${N2},${N3}\]" comdata.f90 || { echo "nP substitution failed" >&2; exit 4; }
grep -q "maxstep	=${MAXSTEP}" comdata.f90 || { echo "maxstep substitution failed" >&2; exit 4; }

# bottom-wall block of a main: from the first "if(bottom) then" after the
# "! Boundary Conditions" marker up to the "if(top) then" that follows it
if [ -n "$SPLICE" ]; then
  awk '/! Boundary Conditions/{f=1} f==1 && /if\(bottom\) then/{f=2} f==2 && /if\(top\) then/{exit} f==2{print}' \
      "$SPLICE" > bc_study.f90
  [ -s bc_study.f90 ] || { echo "no bottom-wall block in $SPLICE" >&2; exit 4; }
  awk -v B=bc_study.f90 '/! Boundary Conditions/{f=1}
      f==1 && /if\(bottom\) then/ { while ((getline l < B) > 0) print l; f=2; next }
      f==2 && /if\(top\) then/ { f=3 }
      f==2 { next }
      { print }' "$MAINSRC" > main_src.f90
  cmp -s main_src.f90 "$MAINSRC" && { echo "splice changed nothing" >&2; exit 4; }
else
  cp "$MAINSRC" main_src.f90
fi

# output cadence (every NWRITE steps, alternating on nstep/100 as v1.9) and
# full-precision solution files
WFMT="ES25.17"
[ -n "$FMTSUF" ] && WFMT="E22.15"
sed -e "s/if(mod(nstep,100).eq.0) then/if(mod(nstep,${NWRITE}).eq.0) then/" \
    -e "s/if(mod(nstep,1000).eq.0) then/if(mod(nstep,${NWRITE}).eq.0) then/" \
    -e "s/if(mod(nstep\/1000,2).eq.0) then/if(mod(nstep\/100,2).eq.0) then/" \
    -e "s/write(10,'(3(1X,F10.6),7(1X,E22.15),1X,I1)')/write(10,'(3(1X,F10.6),7(1X,${WFMT}),1X,I1)')/" \
    main_src.f90 > main_ref.f90
grep -q "mod(nstep,${NWRITE}).eq.0" main_ref.f90 || { echo "cadence substitution failed" >&2; exit 4; }
grep -q "7(1X,${WFMT}),1X,I1" main_ref.f90 || { echo "format substitution failed" >&2; exit 4; }

printf '      module mpi\n      include '"'"'mpif.h'"'"'\n      end module mpi\n' > mpi_mod.f

MCM=""
# static arrays above 2 GiB need the medium code model
if [ $(( N1 * N2 * N3 )) -gt 4000000 ]; then MCM="-mcmodel=medium"; fi
$FC -O2 $MCM -ffixed-form -I"$MPIINC" -c mpi_mod.f -o mpi_mod.o
$FC -O2 $MCM -c comdata.f90 -o comdata.o
$FC -O2 $MCM -c main_ref.f90 -o main_ref.o
$FC -O2 $MCM -o main.x main_ref.o comdata.o mpi_mod.o -L"$MPILIB" -lmpifort -lmpi -Wl,-rpath,"$MPILIB"
echo "$OUT/main.x"
