"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker; the product never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

RHOMIN, UUMIN, GAMMAMAX = 1.0e-4, 1.0e-6, 25.0  # initbase.c:459-471
LOCS_USED = 7  # FACE1, FACE2, FACE3, CORN1, CORN2, CORN3, CENT


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = C.CDLL(LIB)
        d = C.POINTER(C.c_double)
        i32 = C.POINTER(C.c_int32)
        i64 = C.POINTER(C.c_int64)
        _lib.or_c2p_batch.argtypes = [C.c_int64, d, d, d, d, d, C.c_double, d, i32, i32]
        _lib.or_primtoU_batch.argtypes = [C.c_int64, d, d, d, d, d, C.c_double, d]
        _lib.or_c2p_converged_batch.argtypes = [C.c_int64, d, d, d, d, d, C.c_double, d, i32, C.c_int, d, d, d]
        _lib.or_step.argtypes = [C.POINTER(OrGrid), d, d, i64, i64]
        _lib.or_fluxcalc.argtypes = [C.POINTER(OrGrid), d, d, d, d, d]
        _lib.or_bound_prim.argtypes = [C.POINTER(OrGrid), d]
        _lib.or_reconstruct.argtypes = [C.POINTER(OrGrid), d, C.c_int, d, d]
        _lib.or_floors.argtypes = [C.c_int64, d, C.c_double, C.c_double, d]
        _lib.or_post_fixup.argtypes = [C.POINTER(OrGrid), d, i32, C.c_int]
        _lib.or_step_unew.argtypes = [C.POINTER(OrGrid), d, d, d, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        _lib.or_step_fv.argtypes = [C.POINTER(OrGrid), d, d, d, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        _lib.or_cons_from_prims.argtypes = [C.POINTER(OrGrid), d, d]
    return _lib


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


# scheme number of the fixtures ("weno" entry, the golden harness's second
# argument) -> (oracle lim, flux, timeorder, finite volume): MC + LaxF + RK2,
# WENO5BND + HLL + RK2, PARA + HLL + RK4, the last with the finite-volume
# step (the torus file's own choice), MC + LaxF + TVD RK3, MC + LaxF + Euler,
# MC + FORCE, PARA + MUSTA, WENO5BND + HLLLAXF1 (midpoint RK2), WENO5BND
# finite volume, VANL + LaxF, MINM + HLL, DONOR + LaxF (5th entry: the
# three-point limiter, MC when absent)
SCHEMES = {0: (0, 0, 2, 0), 1: (1, 1, 2, 0), 2: (2, 1, 4, 0), 3: (2, 1, 4, 1), 4: (0, 0, 3, 0), 5: (0, 0, 1, 0),
           6: (0, 2, 2, 0), 7: (2, 3, 2, 0), 8: (1, 4, 2, 0), 9: (1, 1, 2, 1),
           10: (0, 0, 2, 0, 1), 11: (0, 1, 2, 0, 2), 12: (0, 0, 2, 0, 0)}


class OrGrid(C.Structure):
    _fields_ = [("n1", C.c_int), ("n2", C.c_int), ("n3", C.c_int), ("nb", C.c_int),
                ("lim", C.c_int), ("flux", C.c_int),
                ("gam", C.c_double), ("cour", C.c_double), ("dx", C.c_double * 4),
                ("rhomin", C.c_double), ("uumin", C.c_double), ("gammamax", C.c_double),
                ("defcon", C.c_double), ("safe", C.c_double),
                ("gcov", C.POINTER(C.c_double)), ("gcon", C.POINTER(C.c_double)),
                ("gcovpert", C.POINTER(C.c_double)), ("gdet", C.POINTER(C.c_double)),
                ("conn", C.POINTER(C.c_double)), ("floors", C.POINTER(C.c_double)),
                ("timeorder", C.c_int), ("fv", C.c_int), ("slim", C.c_int)]


def c2p_batch(gcov, gcon, gdet, gcovpert, U, gam, guess):
    n = U.shape[0]
    arrs = [np.ascontiguousarray(x, dtype=np.float64) for x in (gcov, gcon, gdet, gcovpert, U)]
    pr = np.ascontiguousarray(guess, dtype=np.float64).copy()
    pflag = np.zeros(n, np.int32)
    iters = np.zeros(n, np.int32)
    lib().or_c2p_batch(n, *[_p(a) for a in arrs], gam, _p(pr),
                       pflag.ctypes.data_as(C.POINTER(C.c_int32)),
                       iters.ctypes.data_as(C.POINTER(C.c_int32)))
    return pr, pflag, iters


def c2p_converged(gcov, gcon, gdet, gcovpert, U, gam, guess, start_prims=None):
    """The converged root of the reference's FP64 Noble 2D system in long
    double (or_c2p_root.c): (primitives, reference codes, (W, vsq) of the
    root, scaled residuals at the root).  start_prims: start the long double
    Newton from these primitives instead of the reference's solution."""
    n = U.shape[0]
    arrs = [np.ascontiguousarray(x, dtype=np.float64) for x in (gcov, gcon, gdet, gcovpert, U)]
    pr = np.ascontiguousarray(guess, dtype=np.float64).copy()
    pflag = np.zeros(n, np.int32)
    Wv = np.zeros((n, 2))
    res = np.zeros((n, 2))
    sp = np.ascontiguousarray(start_prims if start_prims is not None else np.zeros((1, 8)), dtype=np.float64)
    lib().or_c2p_converged_batch(n, *[_p(a) for a in arrs], gam, _p(pr), pflag.ctypes.data_as(C.POINTER(C.c_int32)),
                                 1 if start_prims is not None else 0, _p(sp), _p(Wv), _p(res))
    return pr, pflag, Wv, res


def primtoU_batch(gcov, gcon, gdet, gcovpert, pr, gam):
    n = pr.shape[0]
    arrs = [np.ascontiguousarray(x, dtype=np.float64) for x in (gcov, gcon, gdet, gcovpert, pr)]
    U = np.zeros((n, 8))
    lib().or_primtoU_batch(n, *[_p(a) for a in arrs], gam, _p(U))
    return U


def floors_from_r(r):
    """set_density_floors rescaletype 0 (fixup.c:1562-1568), with libm pow."""
    r = np.ascontiguousarray(r, dtype=np.float64)
    out = np.zeros(r.shape + (2,))
    lib().or_floors(r.size, _p(r), RHOMIN, UUMIN, _p(out))
    return out


class Grid:
    """A grid with its geometry tables, in the layout of the reference dumps:
    p[i][j][k][pl] with ghosts; geometry [i][j][loc] with loc in
    (FACE1, FACE2, FACE3, CORN1, CORN2, CORN3, CENT)."""

    def __init__(self, n1, n2, n3, nb, lim, flux, gam, cour, dx, gcov, gcon, gcovpert, gdet, conn,
                 r_cent, defcon=1.0, safe=1.3, timeorder=2, fv=0, slim=3):
        self.keep = [np.ascontiguousarray(x, dtype=np.float64)
                     for x in (gcov, gcon, gcovpert, gdet, conn, floors_from_r(r_cent))]
        g = OrGrid()
        g.n1, g.n2, g.n3, g.nb, g.lim, g.flux = n1, n2, n3, nb, lim, flux
        g.gam, g.cour = gam, cour
        for i in range(4):
            g.dx[i] = dx[i]
        g.rhomin, g.uumin, g.gammamax, g.defcon, g.safe = RHOMIN, UUMIN, GAMMAMAX, defcon, safe
        g.gcov, g.gcon, g.gcovpert, g.gdet, g.conn, g.floors = [_p(a) for a in self.keep]
        g.timeorder, g.fv, g.slim = timeorder, fv, slim
        self.g = g
        self.shape = (n1 + 2 * nb * (n1 > 1), n2 + 2 * nb * (n2 > 1), n3 + 2 * nb * (n3 > 1), 8)

    @classmethod
    def from_fixture(cls, d):
        par = d["params"]
        n1, n2, n3, nb = (int(x) for x in par[:4])
        lim, flux, timeorder, fv, slim = (SCHEMES[int(d["weno"])] + (3,))[:5]
        dx = [0.0, par[12], par[13], par[14]]
        r = d["xv_cent"][:, :, 1, 1]
        return cls(n1, n2, n3, nb, lim, flux, par[5], par[10], dx, d["gcov"], d["gcon"],
                   d["gcovpert"], d["gdet"], d["conn"], r, timeorder=timeorder, fv=fv, slim=slim)

    def step(self, p, dt):
        p = np.ascontiguousarray(p, dtype=np.float64).copy()
        dtv = np.array([dt])
        nf = np.zeros(1, np.int64)
        it = np.zeros(1, np.int64)
        rc = lib().or_step(C.byref(self.g), _p(p), _p(dtv), nf.ctypes.data_as(C.POINTER(C.c_int64)),
                           it.ctypes.data_as(C.POINTER(C.c_int64)))
        if rc:
            raise RuntimeError("oracle step failed")
        return p, float(dtv[0]), int(nf[0]), int(it[0])

    def step_unew(self, p, dt):
        """or_step that also returns unew (the last substep's ucum), the
        array restart_write dumps beside p."""
        p = np.ascontiguousarray(p, dtype=np.float64).copy()
        U = np.zeros(self.shape)
        dtv = np.array([dt])
        nf = np.zeros(1, np.int64)
        it = np.zeros(1, np.int64)
        rc = lib().or_step_unew(C.byref(self.g), _p(p), _p(U), _p(dtv), nf.ctypes.data_as(C.POINTER(C.c_int64)),
                                it.ctypes.data_as(C.POINTER(C.c_int64)))
        if rc:
            raise RuntimeError("oracle step failed")
        return p, U, float(dtv[0])

    def cons_from_prims(self, p):
        """init_conservatives: U(p) of the active zones (zero elsewhere)."""
        p = np.ascontiguousarray(p, dtype=np.float64)
        U = np.zeros(self.shape)
        lib().or_cons_from_prims(C.byref(self.g), _p(p), _p(U))
        return U

    def step_fv(self, p, U, dt):
        """One finite-volume step; U is the carried conserved state."""
        p = np.ascontiguousarray(p, dtype=np.float64).copy()
        U = np.ascontiguousarray(U, dtype=np.float64).copy()
        dtv = np.array([dt])
        nf = np.zeros(1, np.int64)
        it = np.zeros(1, np.int64)
        rc = lib().or_step_fv(C.byref(self.g), _p(p), _p(U), _p(dtv), nf.ctypes.data_as(C.POINTER(C.c_int64)),
                              it.ctypes.data_as(C.POINTER(C.c_int64)))
        if rc:
            raise RuntimeError("oracle step failed")
        return p, U, float(dtv[0]), int(nf[0]), int(it[0])

    def fluxcalc(self, p):
        p = np.ascontiguousarray(p, dtype=np.float64)
        F = [np.zeros(self.shape) for _ in range(3)]
        ndt = np.zeros(3)
        rc = lib().or_fluxcalc(C.byref(self.g), _p(p), _p(F[0]), _p(F[1]), _p(F[2]), _p(ndt))
        if rc:
            raise RuntimeError("oracle fluxcalc failed")
        return F, ndt

    def reconstruct(self, p, direction):
        p = np.ascontiguousarray(p, dtype=np.float64)
        pl = np.zeros(self.shape)
        pr = np.zeros(self.shape)
        lib().or_reconstruct(C.byref(self.g), _p(p), direction, _p(pl), _p(pr))
        return pl, pr

    def bound(self, p):
        p = np.ascontiguousarray(p, dtype=np.float64).copy()
        if lib().or_bound_prim(C.byref(self.g), _p(p)):
            raise RuntimeError("oracle bound failed")
        return p

    def post_fixup_sequence(self, p, flags, finalstep):
        """bound_prim, post_fixup, bound_prim (step_ch.c:132-157) on p with
        the inversion codes `flags` [n1][n2][n3] of the active zones; returns
        p and the full flag array [m1][m2][m3]."""
        p = np.ascontiguousarray(p, dtype=np.float64).copy()
        n1, n2, n3 = self.g.n1, self.g.n2, self.g.n3
        m1, m2, m3 = self.shape[:3]
        b1, b2, b3 = (m1 - n1) // 2, (m2 - n2) // 2, (m3 - n3) // 2
        pf = np.zeros((m1, m2, m3), np.int32)
        pf[b1:b1 + n1, b2:b2 + n2, b3:b3 + n3] = flags
        L = lib()
        if L.or_bound_prim(C.byref(self.g), _p(p)):
            raise RuntimeError("oracle bound failed")
        if L.or_post_fixup(C.byref(self.g), _p(p), pf.ctypes.data_as(C.POINTER(C.c_int32)), int(finalstep)):
            raise RuntimeError("oracle post_fixup failed")
        if L.or_bound_prim(C.byref(self.g), _p(p)):
            raise RuntimeError("oracle bound failed")
        return p, pf
