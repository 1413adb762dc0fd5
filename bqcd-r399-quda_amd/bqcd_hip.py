"""Python host side of the gfx950 BQCD hot-path library (ctypes over the C ABI
declared in include/bqcd_hip.h).

The names mirror the reference's Fortran interface: d / d_dag
(d/D21.F90:29), clover_mult_a / clover_mult_ao / clover_mult_b
(clover/clover_mult_*.F90), mtil / mtil_dag (m_tilde.F90), mtdagmt
(mtdagmt.F90), w_mult / w_mult_dag / w_dagger_w (w_mult.F90), cg (cg.F90:51)
and the forces dsf / dsd / dsg.  Host arrays are numpy float64 in the
reference's layouts:
  spinor (volh, 3, 4, 2); gauge (4, 2, stride, 3, 3, 2); clover
  (2, volh, 2, 18, 2); momenta (4, 2, stride, 8).
Device ("_d") calls take torch tensors in the library's SoA layout:
  spinor (12, volh, 2) float64.

There is no CPU fallback: if libbqcd_hip.so is missing or a call fails, an
exception is raised.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbqcd_hip.so")

OP = {"mtil": 0, "mtil_dag": 1, "mtdagmt": 2, "w_mult": 3, "w_mult_dag": 4, "w_dagger_w": 5}
CLOVER = {"a": 0, "ainv": 1, "b": 2}
PREC = {"f64": 64, "f32": 32, "f16": 16, "f16g": 18}

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p


class HipError(RuntimeError):
    pass


class HmcParams(ctypes.Structure):
    """hb_hmc_params of include/bqcd_hip.h"""
    _fields_ = [("beta", ctypes.c_double), ("kappa", ctypes.c_double), ("csw", ctypes.c_double),
                ("rho", ctypes.c_double), ("tau", ctypes.c_double), ("ntau", ctypes.c_int),
                ("m_scale", ctypes.c_int), ("model", ctypes.c_char), ("integrator", ctypes.c_int),
                ("lam", ctypes.c_double), ("rest", ctypes.c_float), ("maxiter", ctypes.c_int),
                ("log", ctypes.c_int), ("precision", ctypes.c_int), ("delta", ctypes.c_double), ("max_inner", ctypes.c_int),
                ("mre_vectors", ctypes.c_int), ("batch", ctypes.c_int), ("verbose", ctypes.c_int)]


class HmcOut(ctypes.Structure):
    """hb_hmc_out of include/bqcd_hip.h"""
    _fields_ = [(n, ctypes.c_double) for n in (
        "exp_dh", "h_old", "h_new", "sg_old", "sg_new", "sp_old", "sp_new", "sd_old", "sd_new",
        "sf1_old", "sf1_new", "sf2_old", "sf2_new", "plaquette")] + \
        [(n, ctypes.c_int) for n in ("accepted", "cg_ncall", "cg_niter_tot", "cg_niter_max",
                                     "n_force_ir", "n_force_uv")] + [("seconds", ctypes.c_double)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


INTEGRATOR = {"leapfrog": 0, "omelyan": 1}


def multi_block_limit(volh, precision="f16"):
    """largest multi-RHS block size whose block fields fit one buffer (host only)"""
    return lib().hb_multi_block_limit(volh, PREC[precision])


def hmc_params(beta, kappa, csw, rho, tau, ntau, m_scale=1, model="C", integrator="leapfrog",
               lam=0.0, rest=1e-20, maxiter=10000, precision="f64", delta=0.1, max_inner=1000,
               mre_vectors=0, batch=False, verbose=False, log=0):
    return HmcParams(beta, kappa, csw, rho, tau, ntau, m_scale, model.encode(), INTEGRATOR[integrator],
                     lam, rest, maxiter, log, PREC[precision], delta, max_inner, mre_vectors, int(batch),
                     int(verbose))


class NoConvergence(HipError):
    pass


_lib = None


def lib():
    global _lib
    if _lib is None:
        # One HIP runtime per process: PyTorch-ROCm carries its own copy of
        # libamdhip64.so.7.  Loaded first, it is the one this library binds to
        # (same soname); loaded second, it would start a second runtime that
        # sees no GPU.  So torch, when installed, is loaded before the library.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise HipError("libbqcd_hip.so not built (run __graft_entry__.build()): %s" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        L.hb_last_error.restype = ctypes.c_char_p
        L.hb_init.argtypes = [ctypes.c_int * 4, ctypes.c_int * 4, ctypes.c_int, ctypes.POINTER(_vp)]
        L.hb_finalize.argtypes = [_vp]
        L.hb_volh.argtypes = [_vp]
        L.hb_load_gauge.argtypes = [_vp, _dp, ctypes.c_int, ctypes.c_int]
        L.hb_save_gauge.argtypes = [_vp, _dp, ctypes.c_int, ctypes.c_int]
        L.hb_load_clover.argtypes = [_vp, _dp, _dp, _dp]
        L.hb_save_clover.argtypes = [_vp, _dp, _dp, _dp]
        L.hb_dslash.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.hb_clover_mult.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.hb_apply.argtypes = [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp, _dp]
        L.hb_cg.argtypes = [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp, _dp,
                            ctypes.c_float, ctypes.c_int, ctypes.c_int, _ip, _dp]
        L.hb_set_stream.argtypes = [_vp, _vp]
        L.hb_dslash_d.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _vp, _vp]
        L.hb_apply_d.argtypes = [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, _vp, _vp]
        L.hb_cg_d.argtypes = [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, _vp, _vp,
                              ctypes.c_float, ctypes.c_int, ctypes.c_int, _ip, _dp]
        L.hb_set_link_recon.argtypes = [_vp, ctypes.c_int]
        L.hb_dsf_dir.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, ctypes.c_double, _dp,
                                 ctypes.POINTER(ctypes.c_int), ctypes.c_int]
        L.hb_clover_mult_field.argtypes = [_vp, ctypes.c_int, _dp, _dp, _dp]
        L.hb_clover_dsf.argtypes = [_vp, ctypes.c_int, _dp, ctypes.c_int, _dp, _dp, ctypes.c_double]
        L.hb_clover_dsd.argtypes = [_vp, ctypes.c_int, _dp, ctypes.c_int, ctypes.c_double]
        L.hb_clover_action_field.argtypes = [_vp, _dp, _dp]
        L.hb_cg_d_noalloc.argtypes = [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, _vp, _vp,
                                      ctypes.c_float, ctypes.c_int, ctypes.c_int, _ip, _dp]
        L.hb_spinor_to_device.argtypes = [_vp, _vp, _dp]
        L.hb_spinor_to_host.argtypes = [_vp, _dp, _vp]
        L.hb_last_timing.argtypes = [_vp, _dp, _ip]
        L.hb_hmc_trajectory.argtypes = [_vp, ctypes.POINTER(HmcParams), _dp, ctypes.c_int, _dp, _dp,
                                        ctypes.c_double, ctypes.POINTER(HmcOut)]
        L.hb_hmc_trajectory_d.argtypes = [_vp, ctypes.POINTER(HmcParams), _vp, _vp, _vp, ctypes.c_double,
                                          ctypes.POINTER(HmcOut)]
        L.hb_hmc_test_d.argtypes = [_vp, ctypes.POINTER(HmcParams), ctypes.c_int, _vp, _vp, _vp,
                                    ctypes.POINTER(HmcOut)]
        L.hb_hmc_test.argtypes = [_vp, ctypes.POINTER(HmcParams), ctypes.c_int, _dp, ctypes.c_int, _dp, _dp,
                                  ctypes.POINTER(HmcOut)]
        L.hb_gaussian_d.argtypes = [_vp, ctypes.c_ulonglong, ctypes.c_double, _vp, ctypes.c_longlong]
        L.hb_conf_write.argtypes = [ctypes.c_int * 4, ctypes.c_int * 4, _dp, ctypes.c_int, ctypes.c_char_p,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_char_p, ctypes.c_int]
        L.hb_conf_read.argtypes = [ctypes.c_char_p, ctypes.c_int * 4, _dp, ctypes.c_int, _dp]
        L.hb_build_clover.argtypes = [_vp, ctypes.c_double]
        L.hb_plaquette.argtypes = [_vp, _dp]
        L.hb_gauge_random.argtypes = [_vp, ctypes.c_ulonglong, ctypes.c_double]
        L.hb_gauge_smear.argtypes = [_vp, ctypes.c_int, ctypes.c_double]
        L.hb_cg_mixed.argtypes = [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp, _dp,
                                  ctypes.c_float, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_double, ctypes.c_int, _ip, _ip, _dp]
        L.hb_cg_multi.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_double,
                                  ctypes.POINTER(_dp), ctypes.POINTER(_dp), ctypes.POINTER(ctypes.c_float),
                                  ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _ip, _dp]
        L.hb_cg_multi_d.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_double,
                                    ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(ctypes.c_float),
                                    ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _ip, _dp]
        L.hb_multi_block_size.argtypes = [_vp]
        L.hb_multi_tile.argtypes = [_vp]
        L.hb_multi_block_limit.argtypes = [ctypes.c_int, ctypes.c_int]
        L.hb_time_dslash_lp.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _vp, _dp]
        L.hb_apply_lp_d.argtypes = [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                    _vp, _vp]
        L.hb_multi_set_tile.argtypes = [_vp, ctypes.c_int]
        L.hb_dsf.argtypes = [_vp, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp, _dp]
        L.hb_dsd.argtypes = [_vp, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.hb_dsg.argtypes = [_vp, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.hb_hmc_u.argtypes = [_vp, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.hb_clover_action.argtypes = [_vp, _dp]
        L.hb_dsf_d.argtypes = [_vp, _vp, ctypes.c_double, ctypes.c_double, ctypes.c_double, _vp, _vp]
        L.hb_dsd_d.argtypes = [_vp, _vp, ctypes.c_double, ctypes.c_double]
        L.hb_dsg_d.argtypes = [_vp, _vp, ctypes.c_double, ctypes.c_double]
        L.hb_hmc_u_d.argtypes = [_vp, _vp, ctypes.c_double, ctypes.c_double]
        L.hb_gauge_backup.argtypes = [_vp, ctypes.c_int]
        L.hb_gauge_restore.argtypes = [_vp, ctypes.c_int]
        L.hb_cg_mixed_d.argtypes = [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, _vp, _vp,
                                    ctypes.c_float, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_double, ctypes.c_int, _ip, _ip, _dp]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        msg = lib().hb_last_error().decode(errors="replace")
        if rc == -4:
            raise NoConvergence(msg)
        raise HipError("bqcd_hip error %d: %s" % (rc, msg))


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], "need C-contiguous float64"
    return a.ctypes.data_as(_dp)


def _tp(t, numel=0):
    """device pointer of a contiguous float64 GPU tensor holding at least numel
    values: the library reads and writes that many through it, so anything
    else is refused here rather than reaching a kernel"""
    if not (getattr(t, "is_cuda", False) and str(t.dtype) == "torch.float64" and t.is_contiguous()):
        raise HipError("expected a contiguous float64 GPU tensor, got %s of shape %s on %s"
                       % (getattr(t, "dtype", type(t).__name__), tuple(getattr(t, "shape", ())),
                          getattr(t, "device", "the host")))
    if t.numel() < numel:
        raise HipError("device tensor of %d values where %d are needed" % (t.numel(), numel))
    return ctypes.c_void_p(t.data_ptr())


def conf_write(L, u, directory, run=1, traj=0, ensemble=1, former=1, beta=0.0, kappa=0.0, csw=0.0,
               h=0.0, plaq_energy=0.0, bc=(1, 1, 1, -1)):
    """conf_write (conf.F90:384): BQCD configuration files of the host gauge
    field u (BQCD layout, physical links); returns the .info path.  Host only."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = ctypes.create_string_buffer(1024)
    _check(lib().hb_conf_write((ctypes.c_int * 4)(*L), (ctypes.c_int * 4)(*bc), _p(u), u.shape[2],
                               directory.encode(), run, traj, ensemble, former, beta, kappa, csw, h,
                               plaq_energy, out, len(out)))
    return out.value.decode()


def conf_read(info_file, L):
    """conf_read (conf.F90:305): the gauge field (BQCD layout) and the header's
    plaquette energy of a BQCD configuration; checksums are verified.  Host only."""
    vh = L[0] * L[1] * L[2] * L[3] // 2
    u = np.zeros((4, 2, vh, 3, 3, 2))
    plaq = ctypes.c_double(0.0)
    _check(lib().hb_conf_read(info_file.encode(), (ctypes.c_int * 4)(*L), _p(u), vh, ctypes.byref(plaq)))
    return u, plaq.value


class Lattice:
    """One lattice on one GPU: geometry, links and clover term resident in HBM."""

    def __init__(self, L, bc=(1, 1, 1, -1), device=0):
        self.L = tuple(int(x) for x in L)
        self.bc = tuple(int(x) for x in bc)
        self._h = _vp()
        _check(lib().hb_init((ctypes.c_int * 4)(*self.L), (ctypes.c_int * 4)(*self.bc), device,
                             ctypes.byref(self._h)))
        self.volh = lib().hb_volh(self._h)

    def close(self):
        if self._h:
            lib().hb_finalize(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- device tensors handed to the library ---------------------------------
    def _s(self, t):
        """one parity's spinor field: (12, volh, 2) float64 on the GPU"""
        return _tp(t, 24 * self.volh)

    def _m(self, t):
        """momenta: (4, 2, 8, volh) float64 on the GPU"""
        return _tp(t, 64 * self.volh)

    # ---- host allocation helpers ------------------------------------------
    def spinor(self):
        return np.zeros((self.volh, 3, 4, 2))

    # ---- fields -------------------------------------------------------------
    def load_gauge(self, u, flipped=False):
        u = np.ascontiguousarray(u, dtype=np.float64)
        stride = u.shape[2]
        _check(lib().hb_load_gauge(self._h, _p(u), stride, 1 if flipped else 0))

    def save_gauge(self, flipped=False, stride=None):
        stride = stride or self.volh
        u = np.zeros((4, 2, stride, 3, 3, 2))
        _check(lib().hb_save_gauge(self._h, _p(u), stride, 1 if flipped else 0))
        return u

    def load_clover(self, a=None, ainv=None, b=None):
        f = [None if x is None else np.ascontiguousarray(x, dtype=np.float64) for x in (a, ainv, b)]
        _check(lib().hb_load_clover(self._h, *[_p(x) for x in f]))

    def save_clover(self):
        a, ai, b = (np.zeros((2, self.volh, 2, 18, 2)) for _ in range(3))
        _check(lib().hb_save_clover(self._h, _p(a), _p(ai), _p(b)))
        return a, ai, b

    def build_clover(self, csw_kappa):
        """clover_init (clover/clover_init.F90:28) on the device from the loaded links"""
        _check(lib().hb_build_clover(self._h, csw_kappa))

    def plaquette(self):
        """<Re Tr U_p>/3 averaged over all plaquettes"""
        s = ctypes.c_double(0.0)
        _check(lib().hb_plaquette(self._h, ctypes.byref(s)))
        return s.value / (3.0 * 6.0 * 2 * self.volh)

    def plaquette_sum(self):
        """sum over all plaquettes of Re Tr U_p"""
        s = ctypes.c_double(0.0)
        _check(lib().hb_plaquette(self._h, ctypes.byref(s)))
        return s.value

    def gauge_random(self, seed, scale=0.0):
        """seeded synthetic links: exp(i scale w.lambda) for scale > 0, Haar SU(3) otherwise"""
        _check(lib().hb_gauge_random(self._h, seed, scale))

    def gauge_smear(self, sweeps, alpha):
        """APE-like sweeps: U <- P_SU3[(1-alpha) U + alpha/6 staples]"""
        _check(lib().hb_gauge_smear(self._h, sweeps, alpha))

    # ---- operators (host arrays) --------------------------------------------
    def d(self, e, inp, dagger=False):
        out = self.spinor()
        _check(lib().hb_dslash(self._h, e, 1 if dagger else 0, _p(out),
                               _p(np.ascontiguousarray(inp, dtype=np.float64))))
        return out

    def d_dag(self, e, inp):
        return self.d(e, inp, dagger=True)

    def clover_mult(self, which, parity, inp):
        out = self.spinor()
        _check(lib().hb_clover_mult(self._h, CLOVER[which], parity, _p(out),
                                    _p(np.ascontiguousarray(inp, dtype=np.float64))))
        return out

    def clover_mult_a(self, parity, inp):
        return self.clover_mult("a", parity, inp)

    def clover_mult_ao(self, parity, inp):
        return self.clover_mult("ainv", parity, inp)

    def clover_mult_b(self, parity, inp):
        return self.clover_mult("b", parity, inp)

    def apply(self, op, inp, kappa, rho=0.0):
        out = self.spinor()
        _check(lib().hb_apply(self._h, OP[op], kappa, rho, _p(out),
                              _p(np.ascontiguousarray(inp, dtype=np.float64))))
        return out

    def mtil(self, inp, kappa):
        return self.apply("mtil", inp, kappa)

    def mtil_dag(self, inp, kappa):
        return self.apply("mtil_dag", inp, kappa)

    def mtdagmt(self, inp, kappa):
        return self.apply("mtdagmt", inp, kappa)

    def w_mult(self, inp, kappa, rho):
        return self.apply("w_mult", inp, kappa, rho)

    def w_mult_dag(self, inp, kappa, rho):
        return self.apply("w_mult_dag", inp, kappa, rho)

    def w_dagger_w(self, inp, kappa, rho):
        return self.apply("w_dagger_w", inp, kappa, rho)

    # ---- solver -------------------------------------------------------------
    def cg(self, op, b, kappa, rho=0.0, rest=1e-20, maxiter=1000, log=0, x0=None):
        """cg(matrix_mult, x, b, ...) of cg.F90:51: returns (x, iterations, true |r|^2)."""
        x = self.spinor() if x0 is None else np.array(x0, dtype=np.float64, copy=True)
        it = ctypes.c_int(0)
        tr = ctypes.c_double(0.0)
        _check(lib().hb_cg(self._h, OP[op], kappa, rho, _p(x),
                           _p(np.ascontiguousarray(b, dtype=np.float64)), ctypes.c_float(rest),
                           maxiter, log, ctypes.byref(it), ctypes.byref(tr)))
        return x, it.value, tr.value

    def cg_mixed(self, op, b, kappa, rho=0.0, rest=1e-20, maxiter=10000, log=0, x0=None,
                 precision="f16", delta=0.1, max_inner=1000):
        """Mixed-precision CG with reliable updates; returns (x, iters, n_reliable, true |r|^2)."""
        x = self.spinor() if x0 is None else np.array(x0, dtype=np.float64, copy=True)
        it, nrel, tr = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_double(0.0)
        _check(lib().hb_cg_mixed(self._h, OP[op], kappa, rho, _p(x),
                                 _p(np.ascontiguousarray(b, dtype=np.float64)), ctypes.c_float(rest),
                                 maxiter, log, PREC[precision], delta, max_inner, ctypes.byref(it),
                                 ctypes.byref(nrel), ctypes.byref(tr)))
        return x, it.value, nrel.value, tr.value

    def cg_multi(self, op, bs, kappas, rhos=None, csw=0.0, rests=None, maxiter=20000, log=0,
                 x0s=None, precision="f16", block=0):
        """Multi-RHS solve; returns (xs, iters, true |r|^2 per RHS, block size used)."""
        n = len(bs)
        rhos = [0.0] * n if rhos is None else rhos
        rests = [1e-20] * n if rests is None else rests
        xs = [self.spinor() if x0s is None else np.array(x0s[k], dtype=np.float64, copy=True)
              for k in range(n)]
        bs = [np.ascontiguousarray(b, dtype=np.float64) for b in bs]
        xp = (_dp * n)(*[_p(x) for x in xs])
        bp = (_dp * n)(*[_p(b) for b in bs])
        it = (ctypes.c_int * n)()
        tr = (ctypes.c_double * n)()
        _check(lib().hb_cg_multi(self._h, OP[op], n, (ctypes.c_double * n)(*kappas),
                                 (ctypes.c_double * n)(*rhos), csw, xp, bp,
                                 (ctypes.c_float * n)(*rests), maxiter, log, PREC[precision], block,
                                 it, tr))
        return xs, list(it), list(tr), lib().hb_multi_block_size(self._h)

    # ---- forces and link update (host arrays) ----------------------------------
    def time_dslash_lp(self, inp, precision="f16", reps=20):
        """average ms of the reduced-precision hopping kernel (A^-1 D, odd output) of
        the mixed solver on the even device spinor inp (12, volh, 2)"""
        ms = ctypes.c_double(0.0)
        _check(lib().hb_time_dslash_lp(self._h, PREC[precision], reps, self._s(inp), ctypes.byref(ms)))
        return ms.value

    def multi_tile(self):
        """workgroup size of the block kernels of the last cg_multi solve"""
        return lib().hb_multi_tile(self._h)

    def set_multi_tile(self, tile):
        """workgroup size (128 or 256) for cg_multi with an explicit block size"""
        _check(lib().hb_multi_set_tile(self._h, tile))

    def gen(self, stride=None):
        return np.zeros((4, 2, stride or self.volh, 8))

    def dsf(self, a, b, step, kappa, csw_kappa, p=None):
        """dsf (dsf.F90:28): p += hopping + clover fermion force for spinors a, b"""
        p = self.gen() if p is None else np.array(p, dtype=np.float64, copy=True)
        _check(lib().hb_dsf(self._h, _p(p), p.shape[2], step, kappa, csw_kappa,
                            _p(np.ascontiguousarray(a, dtype=np.float64)),
                            _p(np.ascontiguousarray(b, dtype=np.float64))))
        return p

    def dsd(self, step, csw_kappa, p=None):
        """dsd (dsd.F90:28): p += clover-determinant force"""
        p = self.gen() if p is None else np.array(p, dtype=np.float64, copy=True)
        _check(lib().hb_dsd(self._h, _p(p), p.shape[2], step, csw_kappa))
        return p

    def dsg(self, step, beta, p=None):
        """dsg (dsg.F90:28): p += Wilson gauge force"""
        p = self.gen() if p is None else np.array(p, dtype=np.float64, copy=True)
        _check(lib().hb_dsg(self._h, _p(p), p.shape[2], step, beta))
        return p

    def hmc_u(self, p, step, csw_kappa=0.0):
        """hmc_u (hmc_u.F90:28): U <- exp(i step p.lambda) U, reunitarised, clover rebuilt"""
        p = np.ascontiguousarray(p, dtype=np.float64)
        _check(lib().hb_hmc_u(self._h, _p(p), p.shape[2], step, csw_kappa))

    def clover_action(self):
        """clover_action (clover/clover_action.F90:29) of the odd sites"""
        s = ctypes.c_double(0.0)
        _check(lib().hb_clover_action(self._h, ctypes.byref(s)))
        return s.value

    def gauge_backup(self, slot=0):
        _check(lib().hb_gauge_backup(self._h, slot))

    def gauge_restore(self, slot=0):
        _check(lib().hb_gauge_restore(self._h, slot))

    # ---- device-resident API (torch tensors, SoA layout) ---------------------
    def set_stream(self, stream=None):
        _check(lib().hb_set_stream(self._h, ctypes.c_void_p(stream.cuda_stream) if stream else None))

    def to_device(self, host_spinor, out):
        _check(lib().hb_spinor_to_device(self._h, self._s(out),
                                         _p(np.ascontiguousarray(host_spinor, dtype=np.float64))))
        return out

    def to_host(self, t):
        h = self.spinor()
        _check(lib().hb_spinor_to_host(self._h, _p(h), self._s(t)))
        return h

    def dslash_d(self, e, dagger, out, inp):
        _check(lib().hb_dslash_d(self._h, e, 1 if dagger else 0, self._s(out), self._s(inp)))

    def apply_d(self, op, kappa, rho, out, inp):
        _check(lib().hb_apply_d(self._h, OP[op], kappa, rho, self._s(out), self._s(inp)))

    def apply_lp_d(self, op, kappa, rho, out, inp, precision="f16"):
        """out = op inp with the mixed solver's reduced-precision operator
        (inp stored at that precision, the result decoded to FP64)"""
        _check(lib().hb_apply_lp_d(self._h, OP[op], kappa, rho, PREC[precision], self._s(out), self._s(inp)))

    def cg_d(self, op, x, b, kappa, rho=0.0, rest=1e-20, maxiter=1000, log=0):
        it = ctypes.c_int(0)
        tr = ctypes.c_double(0.0)
        _check(lib().hb_cg_d(self._h, OP[op], kappa, rho, self._s(x), self._s(b), ctypes.c_float(rest),
                             maxiter, log, ctypes.byref(it), ctypes.byref(tr)))
        return it.value, tr.value

    # ---- entry points of the LIBD / LIBCLOVER plugins (host arrays) ----------
    def dsf_dir(self, mu, forward, p, b, a, s, u, nn):
        """dsf_{x,y,z,t}{f,b} (d/DSFxyzt.F90:77-117): p (volh, 8) of one (parity,
        mu) accumulates in place; b, a (n_tot, 3, 4, 2) spinors with any halo
        sites; u (volh, 3, 3, 2) links of (parity, mu); nn 1-based int32"""
        assert p.flags["C_CONTIGUOUS"] and p.dtype == np.float64 and p.shape == (self.volh, 8)
        b = np.ascontiguousarray(b, dtype=np.float64)
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.shape == b.shape and a.shape[0] >= self.volh
        nn = np.ascontiguousarray(nn, dtype=np.int32)
        assert nn.shape == (self.volh,)
        _check(lib().hb_dsf_dir(self._h, mu, int(forward), _p(p), _p(b), _p(a), s,
                                _p(np.ascontiguousarray(u, dtype=np.float64)),
                                nn.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), a.shape[0]))

    def clover_mult_field(self, mode, clov_parity, inp):
        """clover_mult_a (mode 0) / clover_mult_b (mode 1) with one parity of a
        clover field (volh, 2, 18, 2)"""
        out = self.spinor()
        _check(lib().hb_clover_mult_field(self._h, mode, _p(np.ascontiguousarray(clov_parity)), _p(out),
                                          _p(np.ascontiguousarray(inp))))
        return out

    def clover_dsf(self, eo, p, b, a, s):
        """clover_dsf (clover/clover_d.F90:78-106) on the loaded physical links;
        p (4, 2, volh, 8) accumulates in place"""
        _check(lib().hb_clover_dsf(self._h, eo, _p(p), p.shape[2], _p(np.ascontiguousarray(b)),
                                   _p(np.ascontiguousarray(a)), s))

    def clover_dsd(self, eo, p, s):
        """clover_dsd (clover/clover_d.F90:40-75) with the loaded LDL^+ field"""
        _check(lib().hb_clover_dsd(self._h, eo, _p(p), p.shape[2], s))

    def clover_action_field(self, b_parity):
        """clover_action (clover/clover_action.F90:29-62) of one parity (volh, 2, 18, 2)"""
        s = ctypes.c_double(0.0)
        _check(lib().hb_clover_action_field(self._h, _p(np.ascontiguousarray(b_parity)), ctypes.byref(s)))
        return s.value

    def set_link_recon(self, recon):
        """links the FP64 CG iterates with: 12 (rows 0 and 1, row 2 rebuilt; the
        default, used while the links are SU(3)) or 18"""
        _check(lib().hb_set_link_recon(self._h, recon))

    def cg_d_noalloc(self, op, x, b, kappa, rho=0.0, rest=1e-20, maxiter=1000, log=0):
        """cg_d with the library's static work area: allocates no device memory
        (volh <= HB_NOALLOC_MAX_VOLH)"""
        it = ctypes.c_int(0)
        tr = ctypes.c_double(0.0)
        _check(lib().hb_cg_d_noalloc(self._h, OP[op], kappa, rho, self._s(x), self._s(b),
                                     ctypes.c_float(rest), maxiter, log, ctypes.byref(it), ctypes.byref(tr)))
        return it.value, tr.value

    def cg_mixed_d(self, op, x, b, kappa, rho=0.0, rest=1e-20, maxiter=10000, log=0,
                   precision="f16", delta=0.1, max_inner=1000):
        it, nrel, tr = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_double(0.0)
        _check(lib().hb_cg_mixed_d(self._h, OP[op], kappa, rho, self._s(x), self._s(b), ctypes.c_float(rest),
                                   maxiter, log, PREC[precision], delta, max_inner, ctypes.byref(it),
                                   ctypes.byref(nrel), ctypes.byref(tr)))
        return it.value, nrel.value, tr.value

    def cg_multi_d(self, op, xs, bs, kappas, rhos=None, csw=0.0, rests=None, maxiter=20000,
                   log=0, precision="f16", block=0):
        n = len(bs)
        rhos = [0.0] * n if rhos is None else rhos
        rests = [1e-20] * n if rests is None else rests
        it = (ctypes.c_int * n)()
        tr = (ctypes.c_double * n)()
        _check(lib().hb_cg_multi_d(self._h, OP[op], n, (ctypes.c_double * n)(*kappas),
                                   (ctypes.c_double * n)(*rhos), csw,
                                   (_vp * n)(*[self._s(t).value for t in xs]),
                                   (_vp * n)(*[self._s(t).value for t in bs]),
                                   (ctypes.c_float * n)(*rests), maxiter, log, PREC[precision],
                                   block, it, tr))
        return list(it), list(tr), lib().hb_multi_block_size(self._h)

    def dsf_d(self, p, a, b, step, kappa, csw_kappa):
        """p (4,2,8,volh) float64 device tensor accumulates the fermion force"""
        _check(lib().hb_dsf_d(self._h, self._m(p), step, kappa, csw_kappa, self._s(a), self._s(b)))

    def dsd_d(self, p, step, csw_kappa):
        _check(lib().hb_dsd_d(self._h, self._m(p), step, csw_kappa))

    def dsg_d(self, p, step, beta):
        _check(lib().hb_dsg_d(self._h, self._m(p), step, beta))

    def hmc_u_d(self, p, step, csw_kappa=0.0):
        _check(lib().hb_hmc_u_d(self._h, self._m(p), step, csw_kappa))

    # ---- HMC trajectory -------------------------------------------------------
    def hmc_trajectory(self, prm, p, eta1, eta2=None, accept_random=-1.0):
        """hmc() (hmc.F90:28) on host draws in BQCD layouts; returns hb_hmc_out as a dict"""
        p = np.ascontiguousarray(p, dtype=np.float64)
        out = HmcOut()
        _check(lib().hb_hmc_trajectory(self._h, ctypes.byref(prm), _p(p), p.shape[2],
                                       _p(np.ascontiguousarray(eta1, dtype=np.float64)),
                                       None if eta2 is None else _p(np.ascontiguousarray(eta2, dtype=np.float64)),
                                       accept_random, ctypes.byref(out)))
        return out.as_dict()

    def hmc_trajectory_d(self, prm, p, eta1, eta2=None, accept_random=-1.0):
        out = HmcOut()
        _check(lib().hb_hmc_trajectory_d(self._h, ctypes.byref(prm), self._m(p), self._s(eta1),
                                         None if eta2 is None else self._s(eta2), accept_random,
                                         ctypes.byref(out)))
        return out.as_dict()

    def hmc_test_d(self, prm, direction, p, eta1=None, eta2=None):
        """hmc()'s reversibility-test modes (hmc.F90:124-210, hmc_test.F90:28-50):
        direction +1 runs a force-accepted trajectory from the momenta p and the
        noise eta1/eta2; direction -1 runs one with the step negated from the
        momenta p and the pseudofermion fields of the trajectory before it.  Both
        overwrite the device tensor p with the final momenta."""
        out = HmcOut()
        _check(lib().hb_hmc_test_d(self._h, ctypes.byref(prm), direction, self._m(p),
                                   None if eta1 is None else self._s(eta1),
                                   None if eta2 is None else self._s(eta2),
                                   ctypes.byref(out)))
        return out.as_dict()

    def hmc_test(self, prm, direction, p, eta1=None, eta2=None):
        """hmc_test_d on host arrays: p (BQCD momenta layout, float64, contiguous)
        is overwritten with the final momenta"""
        if not (isinstance(p, np.ndarray) and p.dtype == np.float64 and p.flags.c_contiguous):
            raise HipError("hmc_test: p must be a C-contiguous float64 array (it is written back)")
        out = HmcOut()
        _check(lib().hb_hmc_test(self._h, ctypes.byref(prm), direction, _p(p), p.shape[2],
                                 None if eta1 is None else _p(np.ascontiguousarray(eta1, dtype=np.float64)),
                                 None if eta2 is None else _p(np.ascontiguousarray(eta2, dtype=np.float64)),
                                 ctypes.byref(out)))
        return out.as_dict()

    def gaussian_d(self, t, seed, var):
        """fill the float64 device tensor t with N(0, var) numbers"""
        _check(lib().hb_gaussian_d(self._h, seed, var, _tp(t), t.numel()))
        return t

    def last_timing(self):
        ms = ctypes.c_double(0.0)
        n = ctypes.c_int(0)
        _check(lib().hb_last_timing(self._h, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value
