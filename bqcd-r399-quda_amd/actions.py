"""The three actions of BQCD's clover-Wilson HMC as torch.autograd.Functions
over the HIP library, with the hand-written force kernels as their adjoints.

Each action is a function of a Lie-algebra displacement omega of the links,
U(omega) = exp(i omega.lambda) U (the map of hmc_u, hmc_u.F90:45-57), given as
a float64 device tensor in the library's momentum layout (4, 2, 8, volh).
Its gradient is minus the corresponding force routine at step 1:

  FermionAction   S_f   = Re phi^+ (M~^+ M~)^-1 phi   grad = -dsf  (dsf.F90:28-110)
  CloverDetAction S_det = clover_action(B_oo)          grad = -dsd  (dsd.F90:28-49,
                                                                      clover_action.F90:29-62)
  GaugeAction     S_g   = -beta/3 sum_p Re Tr U_p      grad = -dsg  (dsg.F90:28-63)

The gradient is the derivative along left multiplication at U(omega), which
is d/d omega exactly at omega = 0 (the usual use: omega = zeros with
requires_grad, then .backward() gives the HMC force).  The links held by the
Lattice are restored after every call; the clover term is rebuilt from them.

The fermion action solves with the FP64 CG by default (precision="f64"), or
with the mixed-precision CG ("f32", "f16"), whose true residual meets the same
bound.  With recompute=True the solution is not kept between forward and
backward: backward moves the links again and re-solves, trading one solve for
the memory of one momentum field.
"""
import torch

import bqcd_hip

_BACKUP_SLOT = 3


def _moved(omega):
    return bool(torch.count_nonzero(omega).item())


class _Displaced:
    """Context manager: links displaced to exp(i omega.lambda) U, clover term
    rebuilt; the original links (and their clover term) are restored on exit."""

    def __init__(self, lat, omega, csw_kappa):
        self.lat, self.omega, self.ck = lat, omega, csw_kappa
        self.moved = _moved(omega)

    def __enter__(self):
        lat = self.lat
        lat.set_stream(torch.cuda.current_stream())
        if self.moved:
            lat.gauge_backup(_BACKUP_SLOT)
            lat.hmc_u_d(self.omega.contiguous(), 1.0, self.ck)
        elif self.ck != 0.0:
            lat.build_clover(self.ck)
        return self

    def __exit__(self, *exc):
        if self.moved:
            self.lat.gauge_restore(_BACKUP_SLOT)
            if self.ck != 0.0:
                self.lat.build_clover(self.ck)
        return False


def _check_omega(lat, omega):
    if omega.dtype != torch.float64 or not omega.is_cuda or tuple(omega.shape) != (4, 2, 8, lat.volh):
        raise ValueError("omega must be a float64 CUDA tensor of shape (4, 2, 8, %d)" % lat.volh)


def _solve(lat, phi, kappa, rest, maxiter, precision):
    x = torch.zeros_like(phi)
    if precision == "f64":
        it, rtr = lat.cg_d("mtdagmt", x, phi, kappa, 0.0, rest, maxiter, log=1)
    else:
        it, _, rtr = lat.cg_mixed_d("mtdagmt", x, phi, kappa, 0.0, rest, maxiter, log=1,
                                    precision=precision)
    return x


def _fermion_force(lat, x, kappa, csw_kappa):
    y = torch.empty_like(x)
    lat.apply_d("mtil", kappa, 0.0, y, x)
    p = torch.zeros((4, 2, 8, lat.volh), dtype=torch.float64, device=x.device)
    lat.dsf_d(p, x, y, 1.0, kappa, csw_kappa)
    return p


class FermionAction(torch.autograd.Function):
    """S_f(omega) = Re phi^+ (M~^+ M~)^-1 phi at links U(omega).

    apply(omega, lat, phi, kappa, csw_kappa, rest=1e-20, maxiter=20000,
          precision="f64", recompute=False)
    phi: float64 CUDA tensor (12, volh, 2), even sites, library layout."""

    @staticmethod
    def forward(ctx, omega, lat, phi, kappa, csw_kappa, rest=1e-20, maxiter=20000,
                precision="f64", recompute=False):
        _check_omega(lat, omega)
        with _Displaced(lat, omega, csw_kappa):
            x = _solve(lat, phi, kappa, rest, maxiter, precision)
            s = torch.sum(phi * x)
            grad = None if recompute else -_fermion_force(lat, x, kappa, csw_kappa)
        ctx.lat, ctx.args = lat, (kappa, csw_kappa, rest, maxiter, precision)
        if recompute:
            ctx.save_for_backward(omega.detach(), phi)
        else:
            ctx.save_for_backward(grad)
        ctx.recompute = recompute
        return s

    @staticmethod
    def backward(ctx, g):
        kappa, csw_kappa, rest, maxiter, precision = ctx.args
        if ctx.recompute:
            omega, phi = ctx.saved_tensors
            with _Displaced(ctx.lat, omega, csw_kappa):
                x = _solve(ctx.lat, phi, kappa, rest, maxiter, precision)
                grad = -_fermion_force(ctx.lat, x, kappa, csw_kappa)
        else:
            (grad,) = ctx.saved_tensors
        return (g * grad,) + (None,) * 8


class CloverDetAction(torch.autograd.Function):
    """S_det(omega) = clover_action of the odd-site clover term at U(omega)."""

    @staticmethod
    def forward(ctx, omega, lat, csw_kappa):
        _check_omega(lat, omega)
        with _Displaced(lat, omega, csw_kappa):
            s = lat.clover_action()
            p = torch.zeros_like(omega)
            lat.dsd_d(p, 1.0, csw_kappa)
        ctx.save_for_backward(-p)
        return torch.tensor(s, dtype=torch.float64, device=omega.device)

    @staticmethod
    def backward(ctx, g):
        (grad,) = ctx.saved_tensors
        return g * grad, None, None


class GaugeAction(torch.autograd.Function):
    """S_g(omega) = -beta/3 sum over plaquettes of Re Tr U_p at U(omega)."""

    @staticmethod
    def forward(ctx, omega, lat, beta):
        _check_omega(lat, omega)
        with _Displaced(lat, omega, 0.0):
            s = -beta / 3.0 * lat.plaquette_sum()
            p = torch.zeros_like(omega)
            lat.dsg_d(p, 1.0, beta)
        ctx.save_for_backward(-p)
        return torch.tensor(s, dtype=torch.float64, device=omega.device)

    @staticmethod
    def backward(ctx, g):
        (grad,) = ctx.saved_tensors
        return g * grad, None, None


def pseudofermion(lat, eta, kappa):
    """phi = M~^+ eta (two-flavour heat bath, hmc_init_phi.F90) for a Gaussian eta"""
    phi = torch.empty_like(eta)
    lat.set_stream(torch.cuda.current_stream())
    lat.apply_d("mtil_dag", kappa, 0.0, phi, eta)
    return phi


__all__ = ["FermionAction", "CloverDetAction", "GaugeAction", "pseudofermion", "bqcd_hip"]
