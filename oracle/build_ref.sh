#!/usr/bin/env bash
# Builds the reference solver (halo-corrected v1.9 physics) from its own
# sources where they lie under /root/reference, for use as a CHECKER only.
#
#   oracle/build_ref.sh N1 N2 N3 MAXSTEP NWRITE
#
# Output: oracle/_ref/n<N1>x<N2>x<N3>_s<MAXSTEP>_w<NWRITE>/main.x (gitignored).
#
# The reference sets its grid and step count as compile-time parameters
# (v1.9/codes/comdata.f90:4,8-9), so the build substitutes them in a copy
# written under oracle/_ref (never committed).  Two further build-time edits
# instrument the output only: the 3D solution file is written every NWRITE
# steps instead of every 100 (main_v1.9.f90:769) and with ES25.17 instead of
# E22.15 (main_v1.9.f90:788) so that a file holds every bit of the state.
# The numerics are untouched.  MPI is the image's MPICH 3.3.2; its Fortran
# module was built by gfortran and cannot be read by flang, so `use mpi` is
# satisfied by compiling MPICH's own mpif.h into a module (what MPICH's
# mpi module is made from); at one rank every exchange is a self-copy.
set -euo pipefail
N1=${1:?N1}; N2=${2:?N2}; N3=${3:?N3}; MAXSTEP=${4:?MAXSTEP}; NWRITE=${5:-100}
REF=/root/reference
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref/n${N1}x${N2}x${N3}_s${MAXSTEP}_w${NWRITE}"
FC=${FC:-/opt/rocm/lib/llvm/bin/flang}
MPIINC=/opt/conda/include
MPILIB=/opt/conda/lib
[ -d "$REF" ] || { echo "reference sources not present: $REF" >&2; exit 3; }
mkdir -p "$OUT"
cd "$OUT"

sed -e "s/=\[130,130,160\]/=[${N1},${N2},${N3}]/" \
    -e "s/maxstep\t=int(5E7)/maxstep\t=${MAXSTEP}/" \
    "$REF/v1.9/codes/comdata.f90" > comdata.f90
grep -q "\[${N1},${N2},${N3}\]" comdata.f90 || { echo "nP substitution failed" >&2; exit 4; }
grep -q "maxstep	=${MAXSTEP}" comdata.f90 || { echo "maxstep substitution failed" >&2; exit 4; }

sed -e "s/if(mod(nstep,100).eq.0) then/if(mod(nstep,${NWRITE}).eq.0) then/" \
    -e "s/write(10,'(3(1X,F10.6),7(1X,E22.15),1X,I1)')/write(10,'(3(1X,F10.6),7(1X,ES25.17),1X,I1)')/" \
    "$REF/v1.10/codes/main_v1.9.f90" > main_ref.f90
grep -q "mod(nstep,${NWRITE}).eq.0" main_ref.f90 || { echo "cadence substitution failed" >&2; exit 4; }
grep -q "ES25.17" main_ref.f90 || { echo "format substitution failed" >&2; exit 4; }

printf '      module mpi\n      include '"'"'mpif.h'"'"'\n      end module mpi\n' > mpi_mod.f

MCM=""
# static arrays above 2 GiB need the medium code model
if [ $(( N1 * N2 * N3 )) -gt 4000000 ]; then MCM="-mcmodel=medium"; fi
$FC -O2 $MCM -ffixed-form -I"$MPIINC" -c mpi_mod.f -o mpi_mod.o
$FC -O2 $MCM -c comdata.f90 -o comdata.o
$FC -O2 $MCM -c main_ref.f90 -o main_ref.o
$FC -O2 $MCM -o main.x main_ref.o comdata.o mpi_mod.o -L"$MPILIB" -lmpifort -lmpi -Wl,-rpath,"$MPILIB"
echo "$OUT/main.x"
