#!/usr/bin/env bash
# Builds the reference's own main program (src/moldin.F90 with its io_m,
# analysis_m and metropolis_m) over the drop-in: the reference's state modules
# (constants_m, utilityfns_m, debug_m, random_m, params_m, timers_m,
# particles_m, system_m, matrix_m, parrinellorahman_m) are compiled from the
# reference tree, the hot-path modules (potential_m, linkcell_m,
# neighbourlist_m, thermostat_m, dynamics_m, quench_m) come from this
# directory, and the program links against libmoldy_hip.so.  Nothing from the
# reference is copied: every file is compiled where it lies, except
# src/metropolis_m.F90, whose text is streamed through sed into the compiler
# because four of its PUBLIC statements carry a type (":25-32 integer,public ::
# get_last_decrease" and the like) that amdflang rejects next to the module
# functions of the same name; the sed expression drops only those types.
#
# usage: build_refmain.sh <reference_src_dir> <build_dir>
# Prints the link's undefined symbols, if any, and exits with the link's status.
set -uo pipefail
SRC=$1
B=$2
HERE=$(cd "$(dirname "$0")" && pwd)
LIB=$(cd "$HERE/../lib" && pwd)
FC=${FC:-/opt/rocm/bin/amdflang}
FFLAGS="-O1 -cpp -fopenmp -w"
mkdir -p "$B"
cd "$B" || exit 2
set -e
for m in constants_m utilityfns_m debug_m random_m params_m timers_m particles_m system_m matrix_m \
         parrinellorahman_m; do
  $FC $FFLAGS -c "$SRC/$m.F90" -o "$m.o"
done
for m in mdk_c mdk_device_m hotpath_m; do
  $FC $FFLAGS -c "$HERE/$m.F90" -o "$m.o"
done
$FC $FFLAGS -c "$SRC/io_m.F90" -o io_m.o
$FC $FFLAGS -c "$SRC/analysis_m.F90" -o analysis_m.o
sed -e 's/^\(\s*\)integer,public :: \(get_last_decrease\|get_success\|get_num_alloy\|energy_converged\)\s*$/\1public :: \2/' \
    "$SRC/metropolis_m.F90" | $FC $FFLAGS -x f95-cpp-input -c - -o metropolis_m.o
$FC $FFLAGS -c "$SRC/moldin.F90" -o moldin.o
set +e
$FC $FFLAGS -o moldy_dropin moldin.o constants_m.o utilityfns_m.o debug_m.o random_m.o params_m.o \
    timers_m.o particles_m.o system_m.o matrix_m.o parrinellorahman_m.o mdk_c.o mdk_device_m.o \
    hotpath_m.o io_m.o analysis_m.o metropolis_m.o -L"$LIB" -lmoldy_hip -Wl,-rpath,"$LIB" 2> link.log
rc=$?
grep -o "undefined symbol: [A-Za-z0-9_]*" link.log | sort -u
exit $rc
