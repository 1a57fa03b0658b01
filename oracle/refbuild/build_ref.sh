#!/usr/bin/env bash
# Build the reference (pseudotensor/harm_finalwham) from its own C sources
# where they lie under /root/reference, linked with the golden-vector
# harness in this directory.  Outputs go only to oracle/_ref/.
#
#   oracle/refbuild/build_ref.sh TAG N1 N2 N3
#
# Notes on the recipe (the reference's own makefile targets icc):
#  * the problem is selected by cfg/init.h, found ahead of the reference's
#    init.h because '-I cfg -I- -I /root/reference' disables the
#    current-directory lookup for quote includes;
#  * gcc 11 and clang 22 reject block-scope 'static' function prototypes,
#    which utoprim_2d_final.c (and the u2p_util.c/utoprim_jon_eos.c it
#    includes) use; that one translation unit is compiled with -Dstatic=,
#    which only changes the linkage of its helpers (its one function with
#    static locals, bin_newt_data, is compiled out by DOHISTOGRAM 0);
#  * the other inversions that fail the same way (utoprim_jon.c,
#    utoprim_1d_final.c, utoprim_5d2_final.c, dudp_calc.c, phys.coldgrmhd.c)
#    are not linked; the four symbols they would provide are only reached
#    for UTOPRIMVERSION values other than UTOPRIM2DFINAL and are left
#    unresolved at link time;
#  * step_ch.c:589 calls bound_prim_user_after_mpi(), which the torus
#    boundary file bounds.fishmon.c does not define; the reference's default
#    boundary file bounds.c defines it (bounds.c:1111, a no-op).  bounds.c is
#    therefore compiled as well, with its three symbols that bounds.fishmon.c
#    also defines renamed by -D so that the torus versions stay in use;
#  * -O2, FP64, SSE2, no FMA contraction: the x86-64 gcc build the fixtures
#    are pinned to.
set -euo pipefail
REF=${HARM_REFERENCE:-/root/reference}
HERE="$(cd "$(dirname "$0")" && pwd)"
OUT="$HERE/../_ref"
TAG=${1:?tag}
N1=${2:?N1}
N2=${3:?N2}
N3=${4:?N3}
BUILD="$OUT/build_$TAG"
mkdir -p "$BUILD"

CFLAGS="-O2 -fcommon -w -ffp-contract=off -mcmodel=medium -DHR_N1=$N1 -DHR_N2=$N2 -DHR_N3=$N3"
INC="-I$HERE/cfg -I- -I$REF"

TUS="initbase restart set_arrays set_grid advance reconstructeno flux wavespeeds boundsflux
fluxct fluxcompute boundsint diag dump dumpgen dump_ener image fail fixup mpi_init mpi_fileio
boundmpi boundmpiint interppoint interpline monointerp jmonointerp smonointerp metric coord
phys step_ch phys.ffde eos vchar transforms sources gaussj lubksb ludcmp mnewt nrutil ranc
tensor utoprimgen dudp_calc_3vel utoprim utoprim_2d utoprim_1d utoprim_1d_opt utoprim_ldz
init.fishmon bounds.fishmon freespace"

pids=()
objs=()
for tu in $TUS; do
  obj="$BUILD/$tu.o"
  objs+=("$obj")
  if [ ! -f "$obj" ] || [ "$REF/$tu.c" -nt "$obj" ] || [ "$HERE/cfg/init.h" -nt "$obj" ]; then
    gcc $CFLAGS $INC -c "$REF/$tu.c" -o "$obj" 2>/dev/null &
    pids+=($!)
  fi
  if [ ${#pids[@]} -ge 8 ]; then wait "${pids[0]}"; pids=("${pids[@]:1}"); fi
done
obj="$BUILD/utoprim_2d_final.o"
objs+=("$obj")
gcc $CFLAGS -Dstatic= $INC -c "$REF/utoprim_2d_final.c" -o "$obj" 2>/dev/null &
pids+=($!)
obj="$BUILD/bounds_default.o"
objs+=("$obj")
gcc $CFLAGS -Dbound_prim_user=refdefault_bound_prim_user -Dapply_bc_line=refdefault_apply_bc_line \
    -Dflip_y=refdefault_flip_y $INC -c "$REF/bounds.c" -o "$obj" 2>/dev/null &
pids+=($!)
for p in "${pids[@]}"; do wait "$p"; done

gcc $CFLAGS $INC -c "$HERE/ref_harness.c" -o "$BUILD/ref_harness.o" 2>/dev/null
gcc -o "$OUT/harness_$TAG" "$BUILD/ref_harness.o" "${objs[@]}" -lm \
    -Wl,--unresolved-symbols=ignore-all -mcmodel=medium
echo "built $OUT/harness_$TAG"
