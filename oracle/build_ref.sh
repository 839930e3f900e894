#!/usr/bin/env bash
# Builds the REFERENCE Fortran hot path (gjackland/moldy, read where it lies
# under /root/reference/src) plus our own driver oracle/ref_driver.F90 into
# oracle/_ref/.  Nothing from the reference is copied: every file is compiled
# in place, except src/thermostat_m.F90, whose text is streamed through sed
# into the compiler's stdin because two of its declarations do not compile
# with amdflang (":54 real(kind_wp), public :: get_snhv" and ":56 ... get_temp"
# conflict with the module functions of the same name).  The sed expression
# drops only the type-spec of those two PUBLIC statements.
#
# Three executables are produced:
#   oracle/_ref/ref_driver      serial, no FMA: golden vectors (tests/golden/)
#   oracle/_ref/ref_driver_omp  -fopenmp: the reference's own OpenMP hot path,
#                               timed as the CPU baseline
#   oracle/_ref/ref_driver_fec  as ref_driver with the iron_carbon material
# usage: oracle/build_ref.sh [reference_src_dir]
set -euo pipefail
SRC=${1:-/root/reference/src}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
FC=${FC:-/opt/rocm/bin/amdflang}
if [ ! -d "$SRC" ]; then
  echo "build_ref: reference sources not found at $SRC (skipped)" >&2
  exit 3
fi
MODS_COMMON1="constants_m utilityfns_m params_m particles_m system_m matrix_m parrinellorahman_m
      random_m lookup_m"
MODS_COMMON2="potential_m linkcell_m neighbourlist_m thermostat_m
      debug_m timers_m dynamics_m analysis_m quench_m"

build_variant() {
  local name=$1; shift
  local material=$1; shift
  local flags="$*"
  local MODS="$MODS_COMMON1 $material $MODS_COMMON2"
  local B="$OUT/build_$name"
  rm -rf "$B"; mkdir -p "$B"
  local objs=""
  for m in $MODS; do
    if [ "$m" = thermostat_m ]; then
      sed -e 's/^\(\s*\)real(kind_wp), public :: get_snhv/\1public :: get_snhv/' \
          -e 's/^\(\s*\)real(kind_wp), public :: get_temp/\1public :: get_temp/' \
          "$SRC/thermostat_m.F90" |
        (cd "$B" && $FC $flags -x f95-cpp-input -c - -o "$B/$m.o" 2> "$B/$m.log")
    else
      (cd "$B" && $FC $flags -c "$SRC/$m.F90" -o "$B/$m.o" 2> "$B/$m.log")
    fi
    objs="$objs $B/$m.o"
  done
  (cd "$B" && $FC $flags -c "$HERE/ref_driver.F90" -o "$B/ref_driver.o")
  $FC $flags -o "$OUT/$name" "$B/ref_driver.o" $objs
}

mkdir -p "$OUT"
# -ffp-contract=off: keep the x86-64 reference free of FMA so that golden
# vectors are reproducible by an operation-order restatement.
build_variant ref_driver fetransmetal -O2 -cpp -DFETRANSMETAL -ffp-contract=off
build_variant ref_driver_omp fetransmetal -O2 -cpp -DFETRANSMETAL -fopenmp
# iron_carbon material (src/iron_carbon.F90), serial and without FMA like ref_driver
build_variant ref_driver_fec iron_carbon -O2 -cpp -DIRON_CARBON -ffp-contract=off
echo "build_ref: built $OUT/ref_driver, $OUT/ref_driver_omp and $OUT/ref_driver_fec"
