"""ctypes binding of libharmgpu.so (include/harmgpu.h).

The library must have been built in-tree (harm_finalwham_amd/build.py);
loading fails loudly otherwise: there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libharmgpu.so")

HG_LIM_DONOR, HG_LIM_VANL, HG_LIM_MINM, HG_LIM_MC, HG_LIM_PARA, HG_LIM_WENO5BND = 0, 1, 2, 3, 4, 17
HG_NOENOFLUX, HG_ENOFINITEVOLUME = 0, 3
HG_FLUX_FORCE, HG_FLUX_MUSTA, HG_FLUX_HLLLAXF1 = 2, 3, 4
HG_FLUX_LAXF, HG_FLUX_HLL = 0, 1
HG_C2P_FP64, HG_C2P_MIXED, HG_C2P_REFINED, HG_C2P_LEARNED = 0, 1, 2, 3

# every symbol include/harmgpu.h declares
EXPORTS = ["hg_ctx_create", "hg_ctx_destroy", "hg_upload_geometry", "hg_mks_geometry_tables",
           "hg_build_mks_geometry", "hg_get_config", "hg_init_torus", "hg_set_prims", "hg_get_prims", "hg_bound_prim", "hg_fluxcalc",
           "hg_step", "hg_step_async", "hg_set_dt", "hg_get_dt", "hg_get_counters",
           "hg_utoprim_batch", "hg_make_c2p_batch", "hg_last_error",
           "hg_profile", "hg_profile_read", "hg_profile_class_name", "hg_c2p_fallbacks",
           "hg_graph_create", "hg_graph_create_grouped", "hg_graph_launch", "hg_graph_destroy", "hg_get_tune", "hg_set_tune", "hg_autotune",
           "hg_mlp_create", "hg_mlp_destroy", "hg_mlp_nparams", "hg_mlp_get_params", "hg_mlp_set_params",
           "hg_mlp_features", "hg_mlp_forward", "hg_mlp_gradient", "hg_mlp_train",
           "hg_utoprim_batch_learned", "hg_post_fixup", "hg_get_pflag", "hg_diagnostics", "hg_diag_record", "hg_rdump_parse",
           "hg_rdump_read", "hg_set_conserved", "hg_get_conserved", "hg_rdump_write", "hg_advance_substep",
           "hg_dump_inversions", "hg_utoprim_batch_noalloc", "hg_set_inversion", "hg_check_division"]
HG_PROF_NCLASS = 11


class HarmError(RuntimeError):
    pass


class HgTune(C.Structure):
    _fields_ = [("bs_flux", C.c_int), ("bs_emf", C.c_int), ("bs_update", C.c_int), ("bs_invert", C.c_int),
                ("weno_xwaves", C.c_int), ("weno_nseg", C.c_int), ("weno_nseg3", C.c_int),
                ("weno_xmarch", C.c_int), ("concurrent", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class HgDiag(C.Structure):
    _fields_ = [("cons", C.c_double * 8), ("divbmax", C.c_double), ("divbavg", C.c_double),
                ("divbmax_ijk", C.c_int32 * 3), ("pdot", C.c_double * 8 * 6),
                ("pdotterms", C.c_double * 8 * 7 * 6), ("pdottermsjet2", C.c_double * 8 * 7 * 6),
                ("pcum", C.c_double * 8 * 6)]

    def as_dict(self):
        import numpy as np
        return {"cons": np.array(self.cons[:]), "divbmax": self.divbmax, "divbavg": self.divbavg,
                "divbmax_ijk": tuple(self.divbmax_ijk[:]),
                "pdot": np.ctypeslib.as_array(self.pdot).copy(),
                "pdotterms": np.ctypeslib.as_array(self.pdotterms).copy(),
                "pdottermsjet2": np.ctypeslib.as_array(self.pdottermsjet2).copy(),
                "pcum": np.ctypeslib.as_array(self.pcum).copy()}


class HgRdumpInfo(C.Structure):
    _fields_ = [("n", C.c_int * 3), ("t", C.c_double), ("tf", C.c_double), ("nstep", C.c_long),
                ("a", C.c_double), ("gam", C.c_double), ("cour", C.c_double), ("DT", C.c_double * 4),
                ("DTr", C.c_long), ("DTdebug", C.c_double), ("counts", C.c_long * 5), ("dt", C.c_double),
                ("lim", C.c_int), ("timeorder", C.c_int), ("fluxmethod", C.c_int), ("fluxb", C.c_int),
                ("utoprimversion", C.c_int), ("failed", C.c_int), ("R0", C.c_double), ("Rin", C.c_double),
                ("Rout", C.c_double), ("hslope", C.c_double), ("bctype", C.c_int * 6), ("realnstep", C.c_long),
                ("defcon", C.c_double), ("safe", C.c_double), ("rhomin", C.c_double), ("uumin", C.c_double),
                ("gammamax", C.c_double), ("gammafail", C.c_double)]

    def as_dict(self):
        out = {}
        for k, _ in self._fields_:
            v = getattr(self, k)
            out[k] = tuple(v[:]) if hasattr(v, "_length_") else v
        return out


class HgConfig(C.Structure):
    _fields_ = [("n1", C.c_int), ("n2", C.c_int), ("n3", C.c_int), ("nb", C.c_int),
                ("lim", C.c_int), ("fluxmethod", C.c_int),
                ("gam", C.c_double), ("cour", C.c_double), ("safe", C.c_double),
                ("defcon", C.c_double), ("dx", C.c_double * 4),
                ("rhomin", C.c_double), ("uumin", C.c_double), ("gammamax", C.c_double),
                ("timeorder", C.c_int), ("doenoflux", C.c_int)]


_lib = None


def lib():
    """Load libharmgpu.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise HarmError("libharmgpu.so not built: run python -m harm_finalwham_amd.build")
    # this binding hands torch device tensors to the library: torch's HIP
    # runtime must come up before the library's first HIP call, or torch
    # finds no device afterwards
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.init()
    except ImportError:
        pass
    L = C.CDLL(LIBPATH)
    d = C.POINTER(C.c_double)
    vp = C.c_void_p
    i64p = C.POINTER(C.c_int64)
    L.hg_ctx_create.argtypes = [C.POINTER(HgConfig), C.POINTER(vp)]
    L.hg_ctx_destroy.argtypes = [vp]
    L.hg_upload_geometry.argtypes = [vp, d, d, d, d, d, d]
    L.hg_mks_geometry_tables.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                         C.c_double, C.c_double, C.c_double, d, d, d, d, d, d, d]
    L.hg_build_mks_geometry.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double]
    L.hg_get_config.argtypes = [vp, C.POINTER(HgConfig)]
    L.hg_init_torus.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                C.c_uint64, d]
    L.hg_set_prims.argtypes = [vp, d]
    L.hg_get_prims.argtypes = [vp, d]
    L.hg_set_conserved.argtypes = [vp, d]
    L.hg_get_conserved.argtypes = [vp, d]
    L.hg_bound_prim.argtypes = [vp, vp]
    L.hg_fluxcalc.argtypes = [vp, d, d, d, d, vp]
    L.hg_step.argtypes = [vp, d, i64p, i64p, vp]
    L.hg_step_async.argtypes = [vp, C.c_int, vp]
    L.hg_set_dt.argtypes = [vp, C.c_double]
    L.hg_get_dt.argtypes = [vp, d]
    L.hg_get_counters.argtypes = [vp, i64p, i64p]
    L.hg_utoprim_batch.argtypes = [C.c_int64, C.c_double, vp, vp, vp, vp, vp, vp, vp, vp, C.c_int, vp]
    L.hg_set_inversion.argtypes = [vp, C.c_int, vp]
    L.hg_utoprim_batch_noalloc.argtypes = [C.c_int64, C.c_double, vp, vp, vp, vp, vp, vp, vp, vp, C.c_int, vp]
    L.hg_dump_inversions.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp]
    L.hg_make_c2p_batch.argtypes = [C.c_int64, C.c_double, C.c_double, C.c_uint64, vp, vp, vp, vp, vp,
                                    vp, vp, vp]
    L.hg_last_error.restype = C.c_char_p
    L.hg_check_division.argtypes = [C.c_int64, C.c_uint64, i64p]
    L.hg_c2p_fallbacks.argtypes = [i64p]
    L.hg_graph_create.argtypes = [C.POINTER(vp), C.c_int, C.c_int, C.POINTER(vp)]
    L.hg_graph_create_grouped.argtypes = [C.POINTER(vp), C.c_int, C.c_int, C.POINTER(vp)]
    L.hg_graph_launch.argtypes = [vp, vp]
    L.hg_graph_destroy.argtypes = [vp]
    L.hg_get_tune.argtypes = [vp, C.POINTER(HgTune)]
    L.hg_set_tune.argtypes = [vp, C.POINTER(HgTune)]
    L.hg_autotune.argtypes = [vp, C.c_int, C.POINTER(HgTune), vp]
    L.hg_profile.argtypes = [vp, C.c_int]
    L.hg_post_fixup.argtypes = [vp, C.POINTER(C.c_int32), C.c_int, vp]
    L.hg_get_pflag.argtypes = [vp, C.POINTER(C.c_int32)]
    L.hg_diagnostics.argtypes = [vp, C.c_double, C.POINTER(HgDiag), vp]
    L.hg_diag_record.argtypes = [vp, C.c_int]
    L.hg_rdump_parse.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, d, d, C.POINTER(HgRdumpInfo)]
    L.hg_rdump_read.argtypes = [vp, C.c_char_p, C.POINTER(HgRdumpInfo), d, vp]
    L.hg_rdump_write.argtypes = [vp, C.c_char_p, C.c_char_p, vp]
    L.hg_advance_substep.argtypes = [vp, C.c_int, C.c_int, C.c_int, d, d, C.c_int, C.c_int, C.c_double, d, vp]
    fp = C.POINTER(C.c_float)
    L.hg_mlp_create.argtypes = [C.c_uint64, C.POINTER(vp)]
    L.hg_mlp_destroy.argtypes = [vp]
    L.hg_mlp_nparams.argtypes = []
    L.hg_mlp_get_params.argtypes = [vp, fp]
    L.hg_mlp_set_params.argtypes = [vp, fp]
    L.hg_mlp_features.argtypes = [C.c_int64, C.c_double, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]
    L.hg_mlp_forward.argtypes = [vp, C.c_int64, vp, vp, vp]
    L.hg_mlp_gradient.argtypes = [vp, C.c_int64, vp, vp, fp, fp, vp]
    L.hg_mlp_train.argtypes = [vp, C.c_int64, vp, vp, C.c_int, C.c_int, C.c_float, fp, vp]
    L.hg_utoprim_batch_learned.argtypes = [vp, C.c_int64, C.c_double, vp, vp, vp, vp, vp, vp, vp, vp, i64p, vp]
    L.hg_profile_read.argtypes = [vp, d, i64p]
    L.hg_profile_class_name.argtypes = [C.c_int]
    L.hg_profile_class_name.restype = C.c_char_p
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise HarmError(lib().hg_last_error().decode())


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Grid:
    """One grid on the GPU (hg_ctx).  Host arrays use the reference layout
    p[i][j][k][pl] with ghost zones."""

    def __init__(self, n1, n2, n3, nb=6, lim=HG_LIM_WENO5BND, fluxmethod=HG_FLUX_HLL, gam=4.0 / 3.0,
                 cour=0.8, dx=(0.0, 1.0, 1.0, 1.0), safe=1.3, defcon=1.0, rhomin=1e-4, uumin=1e-6,
                 gammamax=25.0, timeorder=2, doenoflux=HG_NOENOFLUX):
        cfg = HgConfig()
        cfg.n1, cfg.n2, cfg.n3, cfg.nb, cfg.lim, cfg.fluxmethod = n1, n2, n3, nb, lim, fluxmethod
        cfg.timeorder, cfg.doenoflux = timeorder, doenoflux
        cfg.gam, cfg.cour, cfg.safe, cfg.defcon = gam, cour, safe, defcon
        for k in range(4):
            cfg.dx[k] = dx[k]
        cfg.rhomin, cfg.uumin, cfg.gammamax = rhomin, uumin, gammamax
        self.cfg = cfg
        h = C.c_void_p()
        _check(lib().hg_ctx_create(C.byref(cfg), C.byref(h)))
        self.h = h
        self.n1, self.n2, self.n3 = n1, n2, n3
        nb1, nb2, nb3 = (nb if n > 1 else 0 for n in (n1, n2, n3))
        self.shape = (n1 + 2 * nb1, n2 + 2 * nb2, n3 + 2 * nb3, 8)

    def close(self):
        if self.h:
            lib().hg_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload_geometry(self, gcov, gcon, gcovpert, gdet, conn, r_cent):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (gcov, gcon, gcovpert, gdet, conn, r_cent)]
        _check(lib().hg_upload_geometry(self.h, *[_d(a) for a in arrs]))

    def build_mks_geometry(self, a=0.9375, Rin=1.3, Rout=40.0, R0=0.0, hslope=0.3):
        _check(lib().hg_build_mks_geometry(self.h, a, Rin, Rout, R0, hslope))
        self.cfg = self.config()

    def config(self):
        cfg = HgConfig()
        _check(lib().hg_get_config(self.h, C.byref(cfg)))
        return cfg

    def init_torus(self, rin=6.0, rmax=12.0, kappa=1e-3, beta=100.0, randfact=4e-2, seed=7):
        dt = np.zeros(1)
        _check(lib().hg_init_torus(self.h, rin, rmax, kappa, beta, randfact, seed, _d(dt)))
        return float(dt[0])

    def set_prims(self, p):
        p = np.ascontiguousarray(p, dtype=np.float64)
        assert p.shape == self.shape, (p.shape, self.shape)
        _check(lib().hg_set_prims(self.h, _d(p)))

    def get_prims(self):
        p = np.zeros(self.shape)
        _check(lib().hg_get_prims(self.h, _d(p)))
        return p

    def set_conserved(self, U):
        """The finite-volume scheme's carried conserved state (layout of p)."""
        U = np.ascontiguousarray(U, dtype=np.float64)
        assert U.shape == self.shape, (U.shape, self.shape)
        _check(lib().hg_set_conserved(self.h, _d(U)))

    def get_conserved(self):
        U = np.zeros(self.shape)
        _check(lib().hg_get_conserved(self.h, _d(U)))
        return U

    def bound_prim(self):
        _check(lib().hg_bound_prim(self.h, None))

    def set_inversion(self, mode, mlp=None):
        """The step's inversion (hg_set_inversion): HG_C2P_FP64 (the
        reference's, default), HG_C2P_MIXED, or HG_C2P_LEARNED with the
        network mlp (which must stay alive while the grid uses it)."""
        self._mlp = mlp
        _check(lib().hg_set_inversion(self.h, int(mode), mlp.h if mlp is not None else None))

    def post_fixup(self, codes=None, finalstep=1):
        """bound_prim, post_fixup (fixup.c:37), bound_prim on the current
        primitives; codes: int array [n1][n2][n3] of inversion failure codes
        (None: those of the last inversion)."""
        ptr = None
        if codes is not None:
            codes = np.ascontiguousarray(codes, dtype=np.int32)
            if codes.shape != (self.n1, self.n2, self.n3):
                raise ValueError("codes must have shape (n1, n2, n3)")
            ptr = codes.ctypes.data_as(C.POINTER(C.c_int32))
        _check(lib().hg_post_fixup(self.h, ptr, int(finalstep), None))

    def diagnostics(self, Dt=0.0):
        """constotal, divbmaxavg and diag_flux of the global region (see
        hg_diagnostics); Dt is added to the running flux integral pcum."""
        out = HgDiag()
        _check(lib().hg_diagnostics(self.h, float(Dt), C.byref(out), None))
        return out.as_dict()

    def diag_record(self, on=True):
        """Record diag_flux's boundary terms in every step (hg_diag_record)."""
        _check(lib().hg_diag_record(self.h, int(on)))

    def rdump_read(self, path):
        """Load the primitives of a reference restart dump (then bound_prim);
        returns (unew [n1,n2,n3,8], header dict)."""
        u = np.zeros((self.n1, self.n2, self.n3, 8))
        info = HgRdumpInfo()
        _check(lib().hg_rdump_read(self.h, os.fsencode(path), C.byref(info), _d(u), None))
        return u, info.as_dict()

    def advance_substep(self, pi, pb, pf, CUf, Cunew, timeorder, numtimeorders, dt):
        """One substep of the reference's step loop (hg_advance_substep);
        returns advance's ndt."""
        cuf = np.array(list(CUf) + [0.0] * (4 - len(CUf)), dtype=np.float64)
        cun = np.array(Cunew, dtype=np.float64)
        ndt = np.zeros(1)
        _check(lib().hg_advance_substep(self.h, pi, pb, pf, _d(cuf), _d(cun), timeorder, numtimeorders, float(dt),
                                        _d(ndt), None))
        return float(ndt[0])

    def rdump_write(self, path, header):
        """Write a restart dump in the reference's text format (the header
        line given, then p and unew of every active zone)."""
        _check(lib().hg_rdump_write(self.h, os.fsencode(path), header.encode() if isinstance(header, str) else header,
                                    None))

    def get_pflag(self):
        """Failure codes of the active zones [n1][n2][n3] (FIXED = -1)."""
        out = np.zeros((self.n1, self.n2, self.n3), np.int32)
        _check(lib().hg_get_pflag(self.h, out.ctypes.data_as(C.POINTER(C.c_int32))))
        return out

    def fluxcalc(self):
        F = [np.zeros(self.shape) for _ in range(3)]
        ndt = np.zeros(3)
        _check(lib().hg_fluxcalc(self.h, _d(F[0]), _d(F[1]), _d(F[2]), _d(ndt), None))
        return F, ndt

    def step_ch(self, dt):
        """One RK2 step; returns (next dt, failed inversions, Newton iterations)."""
        dtv = np.array([dt], dtype=np.float64)
        nf = C.c_int64()
        it = C.c_int64()
        _check(lib().hg_step(self.h, _d(dtv), C.byref(nf), C.byref(it), None))
        return float(dtv[0]), nf.value, it.value

    def step_async(self, nsteps, stream=None):
        _check(lib().hg_step_async(self.h, nsteps, stream))

    def set_dt(self, dt):
        _check(lib().hg_set_dt(self.h, dt))

    def profile(self, enable=True):
        _check(lib().hg_profile(self.h, 1 if enable else 0))

    def profile_read(self):
        """{class name: (total ms, intervals)} of the open/last profile."""
        ms = np.zeros(HG_PROF_NCLASS)
        n = np.zeros(HG_PROF_NCLASS, dtype=np.int64)
        _check(lib().hg_profile_read(self.h, _d(ms), n.ctypes.data_as(C.POINTER(C.c_int64))))
        return {lib().hg_profile_class_name(k).decode(): (float(ms[k]), int(n[k])) for k in range(HG_PROF_NCLASS)}

    def tune(self):
        t = HgTune()
        _check(lib().hg_get_tune(self.h, C.byref(t)))
        return t

    def set_tune(self, t):
        _check(lib().hg_set_tune(self.h, C.byref(t)))

    def autotune(self, nsteps=2, stream=None):
        t = HgTune()
        _check(lib().hg_autotune(self.h, nsteps, C.byref(t), stream))
        return t

    def dump_inversions(self):
        """The final substep's inversions of the last step as a batch
        (hg_dump_inversions): a dict of torch float64 tensors in the layout
        of utoprim_batch (U, gcov, gcon, gdet, gcovpert, p_guess)."""
        import torch
        n = self.n1 * self.n2 * self.n3
        f = dict(dtype=torch.float64, device="cuda")
        out = dict(U=torch.empty(8, n, **f), gcov=torch.empty(16, n, **f), gcon=torch.empty(16, n, **f),
                   gdet=torch.empty(n, **f), gcovpert=torch.empty(n, **f), p_guess=torch.empty(8, n, **f))
        ptr = lambda t: C.c_void_p(t.data_ptr())
        _check(lib().hg_dump_inversions(self.h, *[ptr(out[k]) for k in ("U", "gcov", "gcon", "gdet", "gcovpert",
                                                                         "p_guess")], None))
        return out

    def get_dt(self):
        v = np.zeros(1)
        _check(lib().hg_get_dt(self.h, _d(v)))
        return float(v[0])


def rdump_parse(path, n1, n2, n3):
    """Read a restart dump of the reference on the host: (p [n1,n2,n3,8],
    unew [n1,n2,n3,8], header dict)."""
    p = np.zeros((n1, n2, n3, 8))
    u = np.zeros((n1, n2, n3, 8))
    info = HgRdumpInfo()
    _check(lib().hg_rdump_parse(os.fsencode(path), n1, n2, n3, _d(p), _d(u), C.byref(info)))
    return p, u, info.as_dict()


def mks_geometry_tables(n1, n2, n3, nb=6, a=0.9375, Rin=1.3, Rout=40.0, R0=0.0, hslope=0.3):
    """Host geometry tables of the modified Kerr-Schild grid (no GPU needed).
    Returns a dict with gcov/gcon [g1,g2,7,4,4], gcovpert [g1,g2,7,4],
    gdet [g1,g2,7] (slots FACE1 FACE2 FACE3 CORN1 CORN2 CORN3 CENT; g = m+1),
    conn [m1,m2,4,4,4], r_cent [m1,m2] and dx[4]."""
    nbs = [nb if n > 1 else 0 for n in (n1, n2)]
    m1, m2 = n1 + 2 * nbs[0], n2 + 2 * nbs[1]
    g1, g2 = m1 + 1, m2 + 1
    out = dict(gcov=np.zeros((g1, g2, 7, 4, 4)), gcon=np.zeros((g1, g2, 7, 4, 4)),
               gcovpert=np.zeros((g1, g2, 7, 4)), gdet=np.zeros((g1, g2, 7)),
               conn=np.zeros((m1, m2, 4, 4, 4)), r_cent=np.zeros((m1, m2)), dx=np.zeros(4))
    _check(lib().hg_mks_geometry_tables(n1, n2, n3, nb, a, Rin, Rout, R0, hslope,
                                        *[_d(out[k]) for k in ("gcov", "gcon", "gcovpert", "gdet", "conn",
                                                               "r_cent", "dx")]))
    return out


def utoprim_batch(U, gcov, gcon, gdet, gcovpert, pr, pflag, iters=None, gam=4.0 / 3.0, mode=HG_C2P_FP64,
                  stream=None):
    """Utoprimgen over a batch (utoprimgen.c:23) on device tensors in
    structure-of-arrays layout: U[8,n], gcov[16,n], gcon[16,n], gdet[n],
    gcovpert[n] (tt component), pr[8,n] guess in/solution out, pflag[n] int32,
    iters[n] int32 or None.  Arguments are torch CUDA tensors."""
    n = U.shape[1]
    ptr = lambda t: C.c_void_p(t.data_ptr()) if t is not None else None
    _check(lib().hg_utoprim_batch(n, gam, ptr(U), ptr(gcov), ptr(gcon), ptr(gdet), ptr(gcovpert), ptr(pr),
                                  ptr(pflag), ptr(iters), mode, stream))


def utoprim_batch_noalloc(U, gcov, gcon, gdet, gcovpert, pr, pflag, iters=None, gam=4.0 / 3.0,
                          mode=HG_C2P_FP64, stream=None):
    """utoprim_batch through hg_utoprim_batch_noalloc: the same inversion
    with its scratch in the library's static device array (n <= 2^24; calls
    that overlap on one GPU race)."""
    n = U.shape[1]
    ptr = lambda t: C.c_void_p(t.data_ptr()) if t is not None else None
    _check(lib().hg_utoprim_batch_noalloc(n, gam, ptr(U), ptr(gcov), ptr(gcon), ptr(gdet), ptr(gcovpert),
                                          ptr(pr), ptr(pflag), ptr(iters), mode, stream))


def make_c2p_batch(n, gam=4.0 / 3.0, a=0.9375, seed=1, device="cuda"):
    """The batched-inversion benchmark distribution generated on the GPU
    (hg_make_c2p_batch).  Returns a dict of torch float64 tensors in the
    layout of utoprim_batch: U[8,n], gcov[16,n], gcon[16,n], gdet[n],
    gcovpert[n], p_true[8,n], p_guess[8,n]."""
    import torch
    f = dict(dtype=torch.float64, device=device)
    out = dict(U=torch.empty(8, n, **f), gcov=torch.empty(16, n, **f), gcon=torch.empty(16, n, **f),
               gdet=torch.empty(n, **f), gcovpert=torch.empty(n, **f), p_true=torch.empty(8, n, **f),
               p_guess=torch.empty(8, n, **f))
    ptr = lambda t: C.c_void_p(t.data_ptr())
    _check(lib().hg_make_c2p_batch(n, gam, a, seed, *[ptr(out[k]) for k in ("U", "gcov", "gcon", "gdet",
                                                                            "gcovpert", "p_true", "p_guess")],
                                   None))
    return out


def check_division(n, seed=1):
    """hg_check_division: mismatches of the shared-reciprocal division
    against the hardware's x / d over n generated operand sets ([shared
    divisor, reciprocal, divisor 12, sets on the shortcut])."""
    r = np.zeros(4, dtype=np.int64)
    _check(lib().hg_check_division(int(n), int(seed), r.ctypes.data_as(C.POINTER(C.c_int64))))
    return [int(v) for v in r]


def c2p_fallbacks():
    """Zones of the last HG_C2P_MIXED batch inverted by the HG_C2P_REFINED fallback."""
    v = C.c_int64()
    _check(lib().hg_c2p_fallbacks(C.byref(v)))
    return v.value


MLP_NIN, MLP_NH = 8, 64
# offsets into the parameter block (FP32): W1[8][64] b1[64] W2[64][64] b2[64] W3[64] b3 mu[8] sd[8]
MLP_OFF = {"W1": 0, "b1": 512, "W2": 576, "b2": 4672, "W3": 4736, "b3": 4800, "mu": 4801, "sd": 4809}
MLP_NPARAM = 4817


class Mlp:
    """hg_mlp: the learned initial guess of the 2D inversion (8 -> 64 -> 64 -> 1,
    tanh; BF16 weights on the matrix cores, FP32 master copy)."""

    def __init__(self, seed=1):
        h = C.c_void_p()
        _check(lib().hg_mlp_create(seed, C.byref(h)))
        self.h = h
        assert lib().hg_mlp_nparams() == MLP_NPARAM

    def params(self):
        import numpy as np
        w = np.empty(MLP_NPARAM, np.float32)
        _check(lib().hg_mlp_get_params(self.h, w.ctypes.data_as(C.POINTER(C.c_float))))
        return w

    def set_params(self, w):
        import numpy as np
        w = np.ascontiguousarray(w, np.float32)
        if w.shape != (MLP_NPARAM,):
            raise ValueError("parameter block must have %d entries" % MLP_NPARAM)
        _check(lib().hg_mlp_set_params(self.h, w.ctypes.data_as(C.POINTER(C.c_float))))

    def set_normalisation(self, mu, sd):
        w = self.params()
        w[MLP_OFF["mu"]:MLP_OFF["mu"] + MLP_NIN] = mu
        w[MLP_OFF["sd"]:MLP_OFF["sd"] + MLP_NIN] = sd
        self.set_params(w)

    def forward(self, X, stream=None):
        """X: float32 device tensor [n, 8] of raw features -> log(W/D) [n]."""
        import torch
        y = torch.empty(X.shape[0], dtype=torch.float32, device=X.device)
        _check(lib().hg_mlp_forward(self.h, X.shape[0], C.c_void_p(X.data_ptr()), C.c_void_p(y.data_ptr()),
                                    stream))
        return y

    def gradient(self, X, Y, stream=None):
        """Gradient of the mean squared error over (X, Y) w.r.t. the parameter
        block (normalisation entries zero) and the loss; no update."""
        import numpy as np
        g = np.zeros(MLP_NPARAM, np.float32)
        loss = C.c_float()
        _check(lib().hg_mlp_gradient(self.h, X.shape[0], C.c_void_p(X.data_ptr()), C.c_void_p(Y.data_ptr()),
                                     g.ctypes.data_as(C.POINTER(C.c_float)), C.byref(loss), stream))
        return g, loss.value

    def train(self, X, Y, epochs=1, batch=4096, lr=1e-3, stream=None):
        """Adam over the samples in order, minibatches of `batch`; returns the
        mean loss of the last epoch."""
        loss = C.c_float()
        _check(lib().hg_mlp_train(self.h, X.shape[0], C.c_void_p(X.data_ptr()), C.c_void_p(Y.data_ptr()),
                                  epochs, batch, lr, C.byref(loss), stream))
        return loss.value

    def close(self):
        if self.h:
            lib().hg_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def mlp_features(U, gcov, gcon, gdet, p_guess, p_solved=None, gam=4.0 / 3.0, stream=None):
    """Features of each inversion (float32 [n, 8]), with p_solved the targets
    log(W_solved / W_guess) (float32 [n], else None), and a validity mask."""
    import torch
    n = U.shape[1]
    X = torch.empty(n, MLP_NIN, dtype=torch.float32, device=U.device)
    y = torch.empty(n, dtype=torch.float32, device=U.device) if p_solved is not None else None
    ok = torch.empty(n, dtype=torch.int32, device=U.device)
    ptr = lambda t: C.c_void_p(t.data_ptr()) if t is not None else None
    _check(lib().hg_mlp_features(n, gam, ptr(U), ptr(gcov), ptr(gcon), ptr(gdet), ptr(p_guess), ptr(p_solved),
                                 ptr(X), ptr(y), ptr(ok), stream))
    return X, y, ok


def utoprim_batch_learned(mlp, U, gcov, gcon, gdet, gcovpert, pr, pflag, iters=None, gam=4.0 / 3.0,
                          stream=None):
    """utoprim_batch from the better of the learned and the plain start;
    returns (zones solved from the learned start, fallbacks)."""
    n = U.shape[1]
    counts = (C.c_int64 * 2)()
    ptr = lambda t: C.c_void_p(t.data_ptr()) if t is not None else None
    _check(lib().hg_utoprim_batch_learned(mlp.h, n, gam, ptr(U), ptr(gcov), ptr(gcon), ptr(gdet), ptr(gcovpert),
                                          ptr(pr), ptr(pflag), ptr(iters), counts, stream))
    return counts[0], counts[1]


class Graph:
    """hg_graph: nsteps RK2 steps of several grids in one HIP graph; grouped:
    one launch per kernel over all the grids (hg_graph_create_grouped)
    instead of one branch per grid (hg_graph_create)."""

    def __init__(self, grids, nsteps, grouped=False):
        self.h = None
        self.grids = list(grids)  # keep the contexts alive
        arr = (C.c_void_p * len(self.grids))(*[g.h for g in self.grids])
        h = C.c_void_p()
        create = lib().hg_graph_create_grouped if grouped else lib().hg_graph_create
        _check(create(arr, len(self.grids), nsteps, C.byref(h)))
        self.h = h

    def launch(self, stream=None):
        _check(lib().hg_graph_launch(self.h, stream))

    def close(self):
        if self.h:
            lib().hg_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
