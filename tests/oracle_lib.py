"""ctypes wrapper around oracle/liboracle.so (the CPU restatement of the BQCD
hot path).  Test infrastructure: used only by tests, the smoke entry point and
the benchmark's CPU-baseline leg.

All arrays are numpy float64 in BQCD's packed layouts (see oracle/bqcd_oracle.h):
  spinor (volh, 3, 4, 2), gauge (4, 2, volh, 3, 3, 2) [mu, eo, i, col, row, re/im],
  clover (2, volh, 2, 18, 2), generator (4, 2, volh, 8).
"""
import ctypes
import os
import subprocess

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(REPO, "oracle")
LIB_PATH = os.path.join(ORACLE_DIR, "liboracle.so")

_dp = ctypes.POINTER(ctypes.c_double)


class Geom(ctypes.Structure):
    _fields_ = [("L", ctypes.c_int * 4), ("volh", ctypes.c_int), ("nn", ctypes.POINTER(ctypes.c_int))]


class Para(ctypes.Structure):
    _fields_ = [("kappa", ctypes.c_double), ("csw_kappa", ctypes.c_double), ("rho", ctypes.c_double)]


class Op(ctypes.Structure):
    _fields_ = [("g", ctypes.POINTER(Geom)), ("u", _dp), ("a", _dp), ("ainv", _dp), ("p", Para)]


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.or_geom_init.argtypes = [ctypes.POINTER(Geom), ctypes.c_int * 4]
        L.or_geom_free.argtypes = [ctypes.POINTER(Geom)]
        L.or_flip_bc.argtypes = [ctypes.POINTER(Geom), _dp]
        L.or_d.argtypes = [ctypes.POINTER(Geom), ctypes.c_int, ctypes.c_int, _dp, _dp, _dp]
        L.or_clover_mult_a.argtypes = [ctypes.c_int, _dp, _dp, _dp]
        L.or_clover_mult_ao.argtypes = [ctypes.c_int, _dp, _dp]
        L.or_clover_mult_b.argtypes = [ctypes.c_int, _dp, _dp]
        L.or_clover_init.argtypes = [ctypes.POINTER(Geom), _dp, _dp, _dp, _dp, ctypes.c_double]
        L.or_clover_action.argtypes = [ctypes.c_int, _dp]
        L.or_clover_action.restype = ctypes.c_double
        L.or_apply.argtypes = [ctypes.POINTER(Op), ctypes.c_int, _dp, _dp]
        L.or_cg.argtypes = [ctypes.POINTER(Op), ctypes.c_int, _dp, _dp, ctypes.c_float, ctypes.c_int,
                            ctypes.POINTER(ctypes.c_int)]
        L.or_dsf.argtypes = [ctypes.POINTER(Op), _dp, _dp, ctypes.c_double, _dp, _dp, _dp]
        L.or_dsd.argtypes = [ctypes.POINTER(Geom), _dp, _dp, ctypes.c_double, ctypes.c_double, _dp]
        L.or_dsg.argtypes = [ctypes.POINTER(Geom), _dp, _dp, ctypes.c_double, ctypes.c_double]
        L.or_plaquette_sum.argtypes = [ctypes.POINTER(Geom), _dp]
        L.or_plaquette_sum.restype = ctypes.c_double
        L.or_gen2u.argtypes = [_dp, _dp]
        L.or_hmc_u.argtypes = [ctypes.POINTER(Geom), _dp, _dp, ctypes.c_double]
        L.or_ranset.argtypes = [ctypes.c_int64, ctypes.c_int64]
        L.or_ranget.restype = ctypes.c_int64
        L.or_ranf.restype = ctypes.c_double
        L.or_ran_gauss_volh.argtypes = [ctypes.POINTER(Geom), ctypes.c_int, _dp, ctypes.c_double, ctypes.c_int]
        L.or_gauge_exp_random.argtypes = [ctypes.POINTER(Geom), _dp, ctypes.c_int64, ctypes.c_double]
        L.or_spinor_gauss.argtypes = [ctypes.POINTER(Geom), _dp, ctypes.c_int64, ctypes.c_int]
        L.or_momenta_gauss.argtypes = [ctypes.POINTER(Geom), _dp, ctypes.c_int64]
        _lib = L
    return _lib


def _p(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


class Lattice:
    """Geometry plus convenience wrappers around every oracle entry point."""

    def __init__(self, L):
        self.L = tuple(int(x) for x in L)
        self.g = Geom()
        arr = (ctypes.c_int * 4)(*self.L)
        if lib().or_geom_init(ctypes.byref(self.g), arr) != 0:
            raise ValueError("bad lattice %r" % (L,))
        self.volh = self.g.volh

    def __del__(self):
        try:
            lib().or_geom_free(ctypes.byref(self.g))
        except Exception:
            pass

    # --- allocation -------------------------------------------------------
    def spinor(self):
        return np.zeros((self.volh, 3, 4, 2))

    def gauge(self):
        return np.zeros((4, 2, self.volh, 3, 3, 2))

    def clover(self):
        return np.zeros((2, self.volh, 2, 18, 2))

    def gen(self):
        return np.zeros((4, 2, self.volh, 8))

    # --- generators -------------------------------------------------------
    def gauge_exp_random(self, seed, scale=0.3):
        u = self.gauge()
        lib().or_gauge_exp_random(ctypes.byref(self.g), _p(u), seed, scale)
        return u

    def spinor_gauss(self, seed, eo=0):
        s = self.spinor()
        lib().or_spinor_gauss(ctypes.byref(self.g), _p(s), seed, eo)
        return s

    def ran_gauss_volh(self, m, var, eo=0):
        """ran_gauss_volh (ran.F90:28-110) continuing the current ranf stream:
        m complex Gaussians per site of variance var per real component"""
        out = np.zeros((self.volh, m, 2))
        lib().or_ran_gauss_volh(ctypes.byref(self.g), m, _p(out), var, eo)
        return out

    def hmc_draws(self, seed, hasenbusch=True):
        """the random fields of one trajectory of hmc.F90 in its order of draws
        after ranset(seed) (seed None: continuing the current stream, as the
        next trajectory of a run): momenta (hmc_init_p.F90: 8
        ran_gauss_volh(4, ., 1) for mu, eo), then the pseudofermion noise eta1
        and, with Hasenbusch, eta2 (hmc_init_phi.F90: ran_gauss_volh(12, ., 1/2,
        EVEN))"""
        if seed is not None:
            lib().or_ranset(seed, 0)
        p = np.zeros((4, 2, self.volh, 8))
        for mu in range(4):
            for eo in range(2):
                p[mu, eo] = self.ran_gauss_volh(4, 1.0, eo).reshape(self.volh, 8)
        eta1 = self.ran_gauss_volh(12, 0.5).reshape(self.volh, 3, 4, 2)
        eta2 = self.ran_gauss_volh(12, 0.5).reshape(self.volh, 3, 4, 2) if hasenbusch else None
        return p, eta1, eta2

    def momenta_gauss(self, seed):
        p = self.gen()
        lib().or_momenta_gauss(ctypes.byref(self.g), _p(p), seed)
        return p

    # --- operators --------------------------------------------------------
    def flip_bc(self, u):
        u = u.copy()
        lib().or_flip_bc(ctypes.byref(self.g), _p(u))
        return u

    def d(self, e, dagger, inp, u):
        out = self.spinor()
        lib().or_d(ctypes.byref(self.g), e, dagger, _p(out), _p(np.ascontiguousarray(inp)), _p(u))
        return out

    def clover_init(self, u, csw_kappa):
        a, ai, b = self.clover(), self.clover(), self.clover()
        lib().or_clover_init(ctypes.byref(self.g), _p(a), _p(ai), _p(b), _p(u), csw_kappa)
        return a, ai, b

    def clover_mult_a(self, a_par, inp):
        out = self.spinor()
        lib().or_clover_mult_a(self.volh, _p(out), _p(np.ascontiguousarray(a_par)), _p(inp))
        return out

    def clover_mult_b(self, b_par, inp):
        x = inp.copy()
        lib().or_clover_mult_b(self.volh, _p(np.ascontiguousarray(b_par)), _p(x))
        return x

    def op(self, u_flipped, a, ainv, kappa, csw_kappa, rho=0.0):
        o = Op()
        o.g = ctypes.pointer(self.g)
        o.u, o.a, o.ai = _p(u_flipped), _p(a), _p(ainv)
        o.ainv = _p(ainv)
        o.p = Para(kappa, csw_kappa, rho)
        o._keep = (u_flipped, a, ainv)
        return o

    def apply(self, op, which, inp):
        out = self.spinor()
        lib().or_apply(ctypes.byref(op), which, _p(out), _p(np.ascontiguousarray(inp)))
        return out

    def cg(self, op, which, b, rest, maxiter, x0=None):
        x = self.spinor() if x0 is None else x0.copy()
        it = ctypes.c_int(0)
        rc = lib().or_cg(ctypes.byref(op), which, _p(x), _p(np.ascontiguousarray(b)),
                         ctypes.c_float(rest), maxiter, ctypes.byref(it))
        return x, it.value, rc

    def dsf(self, op, u_flipped, step, a, bvec, b_clover):
        p = self.gen()
        u_mut = u_flipped.copy()
        lib().or_dsf(ctypes.byref(op), _p(u_mut), _p(p), step, _p(a), _p(bvec), _p(b_clover))
        return p

    def dsd(self, b_clover, step, csw_kappa, u):
        p = self.gen()
        lib().or_dsd(ctypes.byref(self.g), _p(p), _p(b_clover), step, csw_kappa, _p(u))
        return p

    def dsg(self, u, step, beta):
        p = self.gen()
        lib().or_dsg(ctypes.byref(self.g), _p(p), _p(u), step, beta)
        return p

    def plaquette_sum(self, u):
        return lib().or_plaquette_sum(ctypes.byref(self.g), _p(u))

    def hmc_u(self, u, p, step):
        u = u.copy()
        lib().or_hmc_u(ctypes.byref(self.g), _p(u), _p(np.ascontiguousarray(p)), step)
        return u

    def clover_action(self, b_odd):
        return lib().or_clover_action(self.volh, _p(np.ascontiguousarray(b_odd)))


OPS = {"mtil": 0, "mtil_dag": 1, "mtdagmt": 2, "w_mult": 3, "w_mult_dag": 4, "w_dagger_w": 5}


def gen2u(h):
    u = np.zeros((3, 3, 2))
    lib().or_gen2u(_p(u), _p(np.ascontiguousarray(h, dtype=np.float64)))
    return u  # [col][row][re/im]


def ranset(seed, skip=0):
    lib().or_ranset(seed, skip)


def ranf():
    return lib().or_ranf()


def ranget():
    return lib().or_ranget()
