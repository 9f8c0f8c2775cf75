"""torch.autograd binding of the discrete adjoint of the collapse advance.

    theta : (R,) log-rate multipliers of the network's reactions
            (R = adjoint_ntheta() = 45; M_r = exp(theta_r) scales reaction r
            wherever it enters: the species rates along its stoichiometry, the
            energy rate where the network books its heat, and the Jacobian;
            reaction_table() names them)
    s0    : (18, n) state of each cell at its first advance, field-major:
            xn[0..13] (number densities), T, e, rho, density_driver
    ->  sN: (18, n) state after `nsteps` outer steps (with the density
            perturbations at steps 20, 40, ... when perturb=True)

The forward pass is a recording run on the GPU (the reference integrator on
f_theta, wr_adjoint_record): it checkpoints every outer step and stores the
accepted step sizes.  The backward pass (wr_adjoint_backward) differentiates
that run with its controller decisions frozen -- dt_grid of each step and the
accepted step sizes -- recomputing each step from its checkpoint and the
substeps inside it.  e of s0 does not enter the map (the advance recomputes it
from T), so its gradient is zero.  PyTorch is only the binding: the compute
runs in the HIP library.  Tensors on the GPU stay there: the cells, seeds and
gradients cross the boundary as device addresses (the library copies them on
its own stream after torch's work on them is complete); CPU tensors cross as
float64 host arrays.  Device addresses are only meaningful within one HIP
runtime: torch's wheel carries its own libamdhip64, so torch must be imported
(it maps its runtime) before the library is loaded, which then binds to that
same runtime; the device path checks this and raises otherwise.
"""
from __future__ import annotations

import numpy as np
import torch

from .runtime import CELL_DTYPE, adjoint_ntheta

FIELDS = 18


def cells_from_state(s0):
    """Cells at completed step 0 and time 0 from an (18, n) state array."""
    s0 = np.asarray(s0, dtype=np.float64)
    n = s0.shape[1]
    cells = np.zeros(n, dtype=CELL_DTYPE)
    cells["current"]["xn"] = s0[:14].T
    cells["current"]["T"] = s0[14]
    cells["current"]["e"] = s0[15]
    cells["current"]["rho"] = s0[16]
    cells["density_driver"] = s0[17]
    return cells


# a cell record (216 bytes) as 27 float64 slots: rho, T, e, xn[0..13], time,
# density_driver, (completed_steps, reserved) = 0, seven counters = 0
_SLOTS = 27
_XN = slice(3, 17)


def cells_tensor_from_state(s0):
    """Cell records at completed step 0 and time 0, as an (n, 27) float64
    tensor on s0's device, from an (18, n) state tensor."""
    n = s0.shape[1]
    c = torch.zeros((n, _SLOTS), dtype=torch.float64, device=s0.device)
    c[:, _XN] = s0[:14].T
    c[:, 1] = s0[14]
    c[:, 2] = s0[15]
    c[:, 0] = s0[16]
    c[:, 18] = s0[17]
    return c


def state_of_cells_tensor(c):
    """(18, n) state tensor of an (n, 27) cell-record tensor."""
    return torch.cat([c[:, _XN].T, c[:, 1][None], c[:, 2][None], c[:, 0][None], c[:, 18][None]])


def state_of(cells):
    """(18, n) state array of a cell array."""
    c = cells
    return np.concatenate([c["current"]["xn"].T, c["current"]["T"][None], c["current"]["e"][None],
                           c["current"]["rho"][None], c["density_driver"][None]])


def _hip_runtimes():
    """Distinct HIP runtime libraries mapped into this process (Linux)."""
    try:
        with open("/proc/self/maps") as f:
            return {l.split()[-1] for l in f if "libamdhip64" in l}
    except OSError:
        return set()


def _require_one_hip_runtime():
    if len(_hip_runtimes()) > 1:
        raise RuntimeError(
            "torch and the HIP library run on different HIP runtimes (the library was loaded "
            "before torch): import torch before loading warp-reactor_amd to pass GPU tensors")


class CollapseAdjoint(torch.autograd.Function):
    """sN = collapse(theta, s0) on a wr.Context; gradients by the discrete adjoint."""

    @staticmethod
    def forward(fctx, theta, s0, context, nsteps, hcap, perturb=False):
        if theta.shape != (adjoint_ntheta(),):
            raise ValueError(f"theta must have shape ({adjoint_ntheta()},)")
        M = torch.exp(theta.detach().to(torch.float64)).cpu().numpy()
        fctx.context = context
        fctx.M = M
        fctx.theta_device = theta.device
        if s0.is_cuda:
            _require_one_hip_runtime()
            cells = cells_tensor_from_state(s0.detach().to(torch.float64)).contiguous()
            torch.cuda.synchronize(s0.device)  # the records are complete before the copy
            context.upload_ptr(cells.data_ptr(), cells.shape[0])
            context.adjoint_record(nsteps, M, hcap, perturb)
            context.download_ptr(cells.data_ptr(), cells.shape[0])
            fctx.generation = context.adjoint_generation()  # the tape this forward recorded
            return state_of_cells_tensor(cells)
        cells = cells_from_state(s0.detach().to(torch.float64).numpy())
        context.upload(cells)
        context.adjoint_record(nsteps, M, hcap, perturb)
        context.download(cells)
        fctx.generation = context.adjoint_generation()
        return torch.from_numpy(state_of(cells))

    @staticmethod
    def backward(fctx, gN):
        # fails (WarpReactorError) if a later forward on the same context
        # replaced this forward's tape
        if gN.is_cuda:
            g = gN.detach().to(torch.float64).contiguous()
            n = g.shape[1]
            g0 = torch.empty((FIELDS, n), dtype=torch.float64, device=g.device)
            mbar = torch.empty((adjoint_ntheta(), n), dtype=torch.float64, device=g.device)
            torch.cuda.synchronize(g.device)
            fctx.context.adjoint_backward_ptr(g.data_ptr(), g0.data_ptr(), mbar.data_ptr(),
                                              fctx.generation)
            M = torch.from_numpy(fctx.M).to(device=g.device)
            gtheta = mbar.sum(dim=1) * M  # dM/dtheta = M
            return gtheta.to(fctx.theta_device), g0, None, None, None, None
        g0, mbar = fctx.context.adjoint_backward(gN.detach().to(torch.float64).numpy(),
                                                 generation=fctx.generation)
        gtheta = mbar.sum(axis=1) * fctx.M  # dM/dtheta = M
        return (torch.from_numpy(gtheta).to(fctx.theta_device), torch.from_numpy(g0), None, None,
                None, None)


def collapse(theta, s0, context, nsteps=20, hcap=512, perturb=False):
    """Differentiable collapse advance of the cells in s0 (see module doc)."""
    return CollapseAdjoint.apply(theta, s0, context, nsteps, hcap, perturb)


H2 = 8  # species index of H2 (reproducer.cpp:228-229)
H2_MASS = 3.34706503638e-24


def h2_log_t_loss(sN):
    """Sum over cells of the H2 mass fraction m_H2 n_H2 / rho plus log T."""
    return torch.sum(H2_MASS * sN[H2] / sN[16] + torch.log(sN[14]))
