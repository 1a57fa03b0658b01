"""MI355X implementation of the per-step GRMHD hot path of
pseudotensor/harm_finalwham: HIP kernels for gfx950 behind the C ABI of
include/harmgpu.h, and a thin Python mirror of the reference's entry points
(fluxcalc, bound_prim, step_ch, Utoprimgen) over that ABI."""
from .harmgpu import (HG_C2P_FP64, HG_C2P_MIXED, HG_FLUX_HLL, HG_FLUX_LAXF, HG_LIM_MC,
                      HG_LIM_WENO5BND, HarmError, Grid, lib, utoprim_batch)

__all__ = ["Grid", "HarmError", "lib", "utoprim_batch", "HG_LIM_MC", "HG_LIM_WENO5BND",
           "HG_FLUX_LAXF", "HG_FLUX_HLL", "HG_C2P_FP64", "HG_C2P_MIXED"]
