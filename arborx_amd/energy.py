"""Differentiable neighbour-pair energy over a radius-neighbour CRS.

E = 1/2 sum_i sum_{j in row i, j != i} phi(|x_i - x_j|) with the shifted
Lennard-Jones phi(r) = 4 (r^-12 - r^-6) - phi(cutoff) (no reference
counterpart; the pattern is examples/molecular_dynamics/
example_molecular_dynamics.cpp:128-194).  Forward and backward run in the
HIP library (axb_pair_energy: float pair terms, double sums); the neighbour
set is held fixed.  With recompute_csr=True nothing but the positions is
saved for backward: the CRS is rebuilt there (build + radius query) instead
of being kept alive between forward and backward.
"""
import ctypes

import torch

from . import BoundingVolumeHierarchy, _check, _dev_f32, _ptr, _stream, lib


def pair_energy(xyz, offsets, indices, cutoff, want_energy=True, want_grad=True, stream=None):
    """Returns (E as a 0-d float64 tensor or None, dE/dx (N, 3) float32 or None)."""
    x = _dev_f32(xyz, 3)
    n = x.shape[0]
    e = torch.zeros((), dtype=torch.float64, device="cuda") if want_energy else None
    g = torch.empty((n, 3), dtype=torch.float32, device="cuda") if want_grad else None
    _check(lib().axb_pair_energy(_ptr(x), n, _ptr(offsets), _ptr(indices),
                                 ctypes.c_float(cutoff),
                                 ctypes.c_void_p(e.data_ptr()) if e is not None else None,
                                 _ptr(g) if g is not None else None, _stream(stream)))
    return e, g


def pair_energy_fd(xyz, offsets, indices, cutoff, atoms, h=1e-6, stream=None):
    """FP64 central differences dE/dx for the listed atoms, (m, 3) float64."""
    x = _dev_f32(xyz, 3)
    a = torch.as_tensor(atoms, dtype=torch.int32, device="cuda").contiguous()
    fd = torch.empty((a.numel(), 3), dtype=torch.float64, device="cuda")
    _check(lib().axb_pair_energy_fd(_ptr(x), x.shape[0], _ptr(offsets), _ptr(indices),
                                    ctypes.c_float(cutoff), _ptr(a), a.numel(),
                                    ctypes.c_double(h), _ptr(fd), _stream(stream)))
    return fd


def neighbour_csr(xyz, cutoff, node_format="q16"):
    """Full neighbour list of the points within `cutoff` (row i: every j != i,
    findFullNeighborList semantics) over a tree built from xyz."""
    bvh = BoundingVolumeHierarchy(xyz, node_format=node_format)
    return bvh.neighbor_list(cutoff, full=True)


class PairEnergy(torch.autograd.Function):
    """E(x) with a hand-written backward kernel.

    forward(xyz, cutoff, offsets=None, indices=None, recompute_csr=False):
    if offsets/indices are None the CRS is built from xyz.  The fused
    kernel computes E and dE/dx in one pass over the rows; with
    recompute_csr the CRS and gradient are recomputed in backward instead.
    """

    @staticmethod
    def forward(ctx, xyz, cutoff, offsets=None, indices=None, recompute_csr=False):
        if offsets is None or indices is None:
            offsets, indices = neighbour_csr(xyz.detach(), cutoff)
        if recompute_csr:
            e, _ = pair_energy(xyz.detach(), offsets, indices, cutoff, want_grad=False)
            ctx.save_for_backward(xyz.detach())
            ctx.mode = "recompute"
        else:
            e, g = pair_energy(xyz.detach(), offsets, indices, cutoff, want_grad=True)
            ctx.save_for_backward(g)
            ctx.mode = "saved"
        ctx.cutoff = float(cutoff)
        return e

    @staticmethod
    def backward(ctx, grad_e):
        (t,) = ctx.saved_tensors
        if ctx.mode == "saved":
            g = t
        else:
            offsets, indices = neighbour_csr(t, ctx.cutoff)
            _, g = pair_energy(t, offsets, indices, ctx.cutoff, want_energy=False)
        return (grad_e.to(torch.float32) * g), None, None, None, None
