"""Differentiable cutoff forces for force matching.

``cutoff_forces(force, topology, positions, charges, sigma, epsilon)`` returns
the forces of ``NonbondedCutoffSystemForce`` as a torch tensor whose
backward pass runs the adjoint kernel (``pmnb_adjoint_device``): dL/dx is a
Hessian-vector product and the pairs are recomputed, not stored.  Gradients
reach the raw charges through the scaling by sqrt(332.0636)
(topology/BuildTopology.cpp:92, base/PMConstants.cpp:64) and sigma/epsilon
through the Lorentz-Berthelot table (topology/BuildTopology.cpp:650-661),
both written here in torch so autograd differentiates them.
"""
import torch

from .topology import SQRT_COULOMB_CONSTANT


def lj_table(sigma, epsilon, lj_scaling_factor=1.0):
    """T x T x 4 {A, B, A14, B14}: sigma_ij = (sigma_i + sigma_j) / 2,
    eps_ij = sqrt(eps_i eps_j), A = 4 eps sigma^12, B = 4 eps sigma^6,
    A14 = fudge A, B14 = fudge B (the GROMACS branch of
    BuildTopology.cpp:650-661, as lorentz_berthelot_table)."""
    s = 0.5 * (sigma[:, None] + sigma[None, :])
    e = torch.sqrt(epsilon[:, None] * epsilon[None, :])
    s2 = s * s
    s3 = s2 * s
    s6 = s3 * s3
    s12 = s6 * s6
    A = s12 * e * 4.0
    B = s6 * e * 4.0
    return torch.stack([A, B, lj_scaling_factor * A, lj_scaling_factor * B],
                       dim=-1).contiguous()


class CutoffForceFunction(torch.autograd.Function):
    """forces = F(positions, scaled charges, LJ table) on the device of the
    inputs, in the force's precision.  Backward: the adjoint kernel with the
    upstream gradient; the LJ-table gradient is symmetrised (the table is
    symmetric by construction)."""

    @staticmethod
    def forward(ctx, positions, scaled_charges, table, force, topology):
        stream = torch.cuda.current_stream().cuda_stream
        pos = positions.detach().contiguous()
        q = scaled_charges.detach().contiguous()
        tab = table.detach().contiguous()
        force.set_parameters_device(topology, tab, q, stream)
        frc = torch.zeros_like(pos)
        out = torch.zeros(20, dtype=torch.float64, device=pos.device)
        force.evaluate_device(topology, pos, frc, out, 0, stream=stream)
        ctx.save_for_backward(pos, q, tab)
        ctx.force = force
        ctx.topology = topology
        ctx.energies = out
        return frc

    @staticmethod
    def backward(ctx, grad_forces):
        pos, q, tab = ctx.saved_tensors
        force, topology = ctx.force, ctx.topology
        stream = torch.cuda.current_stream().cuda_stream
        force.set_parameters_device(topology, tab, q, stream)
        g = grad_forces.contiguous()
        gx = torch.zeros_like(pos) if ctx.needs_input_grad[0] else None
        gq = torch.zeros_like(q) if ctx.needs_input_grad[1] else None
        gt = torch.zeros_like(tab) if ctx.needs_input_grad[2] else None
        force.adjoint_device(topology, pos, g, gx, gq, gt, stream=stream)
        return gx, gq, gt, None, None


def cutoff_forces(force, topology, positions, charges, sigma, epsilon,
                  lj_scaling_factor=1.0):
    """Differentiable forces (N x 3, float64) for raw charges (e) and
    per-type sigma/epsilon; every argument may require grad."""
    q = charges * SQRT_COULOMB_CONSTANT
    tab = lj_table(sigma, epsilon, lj_scaling_factor)
    return CutoffForceFunction.apply(positions, q, tab, force, topology)


def force_matching_loss(forces, target):
    """L = sum_i |F_i - F_target,i|^2."""
    d = forces - target
    return (d * d).sum()
