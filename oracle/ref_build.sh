#!/bin/bash
# Builds the reference REWL_Simulator (meese-wj/REWL_Simulator, an MPI program)
# from its own sources where they lie under /root/reference/src, for use as a
# checker of the oracle and as the CPU baseline of kind "reference".  Nothing
# of the reference is kept in this repository: the binary goes to oracle/_ref/
# (git-ignored) and the temporary source copy below is deleted on exit.
#
# The reference is configured at compile time: CMake options become -D macros
# (CMakeLists.txt:31-186; the defaults are used except DIFFERENT_SEEDS, which is
# off so that every walker is seeded with 1, rewl_simulation.hpp:93-102), and
# the lattice size, REWL parameters and field strength are edited in its
# parameter files, as Scripts/msi_setup_Lsize.sh:16 edits SYSTEM_SIZE_L.  Those
# edits need a copy of the source tree; it lives in <output dir>/src_tmp.<name>
# for the duration of the compile only.
#
# With REF_FIELD_SEED set, generate_random_field's clock seed
# (random_fields.hpp:48) is replaced by that constant, the one change to a
# reference source (not a parameter file): it makes a disordered run
# reproducible, which is what rewl_random_field's explicit seed restates.
#
# Usage (environment):
#   REF_NAME      output name: oracle/_ref/rewl_ref_<name>            (required)
#   REF_MODEL     ISING2D | ASHKIN_TELLER2D                           (required)
#   REF_L         SYSTEM_SIZE_L                                       (required)
#   REF_REPLICAS  REWL_Parameters::replicas_per_window                [1]
#   REF_FINAL     REWL_Parameters::final_increment                    [1e-7]
#   REF_SPE       REWL_Parameters::sweeps_per_exchange                [1]
#   REF_SPC       REWL_Parameters::sweeps_per_check                   [1000]
#   REF_FLATNESS  REWL_Parameters::flatness_criterion                 [0.3]
#   REF_ONE_OVER_T ON | OFF                                           [ON]
#   REF_CORR      CORRELATION_LENGTHS ON | OFF                        [ON]
#   REF_DISORDER  none | RFIM | RFAT_BAXTER                           [none]
#   REF_FIELD_H   SYSTEM_FIELD_H (RFIM) or BAXTER_FIELD_H (RFAT)      [0.6 / 0.8]
#   REF_FIELD_SEED constant seed of generate_random_field (disorder only)
#   REF_OPT       optimisation flags       [-O3 -DNDEBUG, CMAKE_BUILD_TYPE Release]
#   REF_OUT_DIR   where the binary goes                               [oracle/_ref]
set -euo pipefail

REF=${REF_ROOT:-/root/reference}/src
HERE=$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)
OUT=${REF_OUT_DIR:-$HERE/_ref}
: "${REF_NAME:?}" "${REF_MODEL:?}" "${REF_L:?}"
REPL=${REF_REPLICAS:-1}
FINAL=${REF_FINAL:-1e-7}
SPE=${REF_SPE:-1}
SPC=${REF_SPC:-1000}
FLAT=${REF_FLATNESS:-0.3}
OPT=${REF_OPT:--O3 -DNDEBUG}
MPI_INC=/opt/conda/include
MPI_LIB=/opt/conda/lib/libmpi.so.12

[ -d "$REF" ] || { echo "reference sources not found at $REF" >&2; exit 2; }
[ -f "$MPI_INC/mpi.h" ] && [ -f "$MPI_LIB" ] || { echo "MPICH not found under /opt/conda" >&2; exit 2; }

mkdir -p "$OUT"
TMP=$OUT/src_tmp.$REF_NAME
rm -rf "$TMP"
trap 'rm -rf "$TMP"' EXIT
cp -r "$REF" "$TMP"
cd "$TMP"

set_line() {   # file, line prefix, replacement line
    grep -q "^$2" "$1" || { echo "no line '$2' in $1" >&2; exit 3; }
    sed -i "s|^$2.*|$3|" "$1"
}

if [ "$REF_MODEL" = ISING2D ]; then
    PARAM=Hamiltonians/Ising2d/ising2d_parameters.cxx
    SRCS="Hamiltonians/Ising2d/source/ising2d_parameter_string.cpp"
else
    PARAM=Hamiltonians/Ashkin_Teller2d/ashkin_teller2d_parameters.cxx
    SRCS="Hamiltonians/Ashkin_Teller2d/source/ashkin_teller2d_parameter_string.cpp"
fi
set_line "$PARAM" "constexpr size_t SYSTEM_SIZE_L = " "constexpr size_t SYSTEM_SIZE_L = $REF_L;"
set_line REWL/rewl_parameters.hpp "    size_t replicas_per_window = " "    size_t replicas_per_window = $REPL;"
set_line REWL/rewl_parameters.hpp "    constexpr LOGDOS_TYPE final_increment = " "    constexpr LOGDOS_TYPE final_increment = $FINAL;"
set_line REWL/rewl_parameters.hpp "    constexpr size_t sweeps_per_exchange = " "    constexpr size_t sweeps_per_exchange = $SPE;"
set_line REWL/rewl_parameters.hpp "    constexpr size_t sweeps_per_check = " "    constexpr size_t sweeps_per_check = $SPC;"
set_line REWL/rewl_parameters.hpp "    constexpr float flatness_criterion = " "    constexpr float flatness_criterion = $FLAT;"

DEFS="-DMPI_ON -DMPICH_SKIP_MPICXX -D$REF_MODEL -DCOLLECT_TIMINGS -DREDUCE_LOGDOS -DTRAPEZOIDAL_RULE -DEQUAL_WINDOWS -DNORMALIZE_BY_STATES"
[ "${REF_ONE_OVER_T:-ON}" = ON ] && DEFS="$DEFS -DONE_OVER_T_ALGORITHM"
[ "${REF_CORR:-ON}" = ON ] && DEFS="$DEFS -DCORRELATION_LENGTHS"
case "${REF_DISORDER:-none}" in
    none) ;;
    RFIM)
        set_line "$PARAM" "constexpr double SYSTEM_FIELD_H = " "constexpr double SYSTEM_FIELD_H = ${REF_FIELD_H:-0.6};"
        DEFS="$DEFS -DRFIM -DRANDOM_DISORDER -DGAUSSIAN_DISORDER";;
    RFAT_BAXTER)
        set_line "$PARAM" "constexpr double BAXTER_FIELD_H = " "constexpr double BAXTER_FIELD_H = ${REF_FIELD_H:-0.8};"
        DEFS="$DEFS -DRFAT_BAXTER -DRANDOM_DISORDER -DGAUSSIAN_DISORDER";;
    *) echo "bad REF_DISORDER" >&2; exit 2;;
esac
if [ -n "${REF_FIELD_SEED:-}" ]; then
    CLOCK='static_cast<std::uint64_t>( std::chrono::high_resolution_clock::now().time_since_epoch().count() );'
    grep -qF "std::uint64_t seed = $CLOCK" Hamiltonians/Disorder/random_fields.hpp || { echo "seed line not found" >&2; exit 3; }
    sed -i "s|std::uint64_t seed = static_cast<std::uint64_t>( std::chrono::high_resolution_clock::now().time_since_epoch().count() );|std::uint64_t seed = ${REF_FIELD_SEED}ull;|" \
        Hamiltonians/Disorder/random_fields.hpp
fi

INCS="-I file_management -I Hamiltonians/grid -I Hamiltonians -I util -I Thermodynamics -I REWL -I $MPI_INC"
COMMON="file_management/file_manager.cpp file_management/file_header.cpp util/command_line_arguments.cpp \
Hamiltonians/grid/grid_setup.cpp Hamiltonians/Address_Books/Square_Periodic_Lattices/square_2d_nearest_neighbors.cpp \
Hamiltonians/Address_Books/Square_Periodic_Lattices/square_2d_nearest_neighbor_functor.cpp REWL/REWL.cpp"
# the reference's own warning flags (CMakeLists.txt:29), except that g++ 11
# (the reference's scripts use 9.2) flags generate_random_field's function
# pointer as maybe-uninitialized; libstdc++ is linked statically so that
# /opt/conda/lib (MPICH's run path) supplies libmpi only
g++ -std=c++17 $OPT -Wall -Werror -pedantic -Wno-error=maybe-uninitialized $DEFS $INCS Simulation/main.cpp $COMMON $SRCS \
    -o "$OUT/rewl_ref_$REF_NAME" -static-libstdc++ "$MPI_LIB" -Wl,-rpath,/opt/conda/lib
echo "$OUT/rewl_ref_$REF_NAME"
