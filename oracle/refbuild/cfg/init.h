/*
 * Problem configuration used when the reference (pseudotensor/harm_finalwham)
 * is compiled from its own sources as a golden-vector generator.
 *
 * The reference selects its problem by editing init.h (dotestone.sh seds
 * TESTNUMBER into it).  This file is found ahead of the reference's init.h
 * on the quote-include path, pulls in the reference's own torus header
 * init.fishmon.h unchanged, and then overrides the handful of compile-time
 * switches that the hot-path configurations need:
 *
 *   N1,N2,N3      grid size, passed as -DHR_N1=.. -DHR_N2=.. -DHR_N3=..
 *   MAXBND 6      ghost width; WENO5BND needs interporder[WENO5BND]/2+1 = 6
 *                 (initbase.c:718 orders_set, initbase.c:792 get_num_bnd_zones_used)
 *   EOMTYPE       EOMGRMHD; init.fishmon.h:170 ships EOMFFDE by mistake
 *
 * Everything else, CHECKSOLUTION 1 (fixup_checksolution, fixup.c:460) and
 * UTOPRIMADJUST UTOPRIMAVG (fixup_utoprim, fixup.c:669) included, is the
 * torus header's own setting.
 */
#include "init.fishmon.h"

#undef N1
#undef N2
#undef N3
#define N1 HR_N1
#define N2 HR_N2
#define N3 HR_N3

#undef MAXBND
#define MAXBND 6

#undef EOMTYPE
#define EOMTYPE EOMGRMHD

