"""Discrete adjoint of the DG RHS and RK steps, for PyTorch.

The adjoint kernels (``nmx_adjoint_setrhs`` / ``nmx_adjoint_evolve``) apply
the exact transpose of the device's forward operator: the scalarwave1 DG RHS
(src/tests/scalarwave1/scalarwave1.c:39-499 with dg.c:30-151 and
numflux1d.c:85-112) and the stages of evolve_RK4_mesh / evolve_sspRK3_mesh /
evolve_Euler_mesh (src/main/evolve/RungeKutta.c:19-188) taken in reverse.
The reference has no adjoint; this module is the harness-side wrapper.

``evolve(mesh, u0, nsteps, every)`` is differentiable with respect to the
initial state u0 (a float64 tensor on the GPU, flat layout u[var*nnodes +
node]).  It returns the states after every ``every`` steps (the last row is
u(nsteps)), keeping device checkpoints at the segment starts.  The backward
pass walks the segments in reverse: it restores the segment's checkpoint,
optionally recomputes the segment forward (``recompute=True``; the adjoint of
the scalarwave1 operator does not read the forward states because the RHS is
linear, so this is off by default), applies the adjoint steps and adds the
incoming gradient of the segment's output.
"""
import torch

from .evolve import REG_ADJ, REG_R, REG_U, Mesh, check, lib  # noqa: F401


def _ptr(t, mesh):
    if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
        raise ValueError("expected a contiguous float64 tensor on the GPU")
    if t.numel() != 5 * mesh.nnodes:
        raise ValueError("a state holds 5*nnodes = %d values, got %d" % (5 * mesh.nnodes, t.numel()))
    return t.data_ptr()


def _handoff(mesh):
    """torch's queued work on the tensors is done before the library reads."""
    torch.cuda.current_stream().synchronize()


def adjoint_setrhs(mesh, lam):
    """L^T lam for the Jacobian L of the DG RHS with respect to u."""
    lam = lam.contiguous()
    out = torch.empty_like(lam)
    _handoff(mesh)
    mesh.copy_in(_ptr(lam, mesh), REG_ADJ)
    check(lib().nmx_adjoint_setrhs(mesh.ctx, REG_R, REG_ADJ))
    mesh.copy_out(_ptr(out, mesh), REG_R)
    mesh.synchronize()
    return out


def adjoint_evolve(mesh, lam, nsteps):
    """(du_n/du_0)^T lam over nsteps steps of evolve_method."""
    lam = lam.contiguous()
    out = torch.empty_like(lam)
    _handoff(mesh)
    mesh.copy_in(_ptr(lam, mesh), REG_ADJ)
    check(lib().nmx_adjoint_evolve(mesh.ctx, int(nsteps)))
    mesh.copy_out(_ptr(out, mesh), REG_ADJ)
    mesh.synchronize()
    return out


def outer_boundary_nodes(mesh):
    """Flat node indices (a LongTensor on the GPU) of the nodes on element
    faces without neighbours, the outer boundary (Node_on_BOUND,
    nmesh_amr_loops.h:125); each node once."""
    import numpy as np
    idx = []
    for e in range(mesh.nelem):
        info = mesh.elm_info(e)
        n, off = int(info[2]), int(info[3])
        ijk = np.arange(n ** 3)
        I = (ijk % n, (ijk // n) % n, ijk // (n * n))
        for f in range(6):
            if info[4 + f] == 0:
                d, side = f // 2, f % 2
                idx.append(off + ijk[I[d] == side * (n - 1)])
    nodes = np.unique(np.concatenate(idx)) if idx else np.zeros(0, dtype=np.int64)
    return torch.from_numpy(nodes.astype(np.int64)).cuda()


def _segments(nsteps, every):
    segs = []
    done = 0
    while done < nsteps:
        k = min(every, nsteps - done)
        segs.append(k)
        done += k
    return segs


class EvolveFunction(torch.autograd.Function):
    """u0 -> [u(k), u(2k), ..., u(nsteps)] with checkpoints every k steps."""

    @staticmethod
    def forward(ctx, u0, mesh, nsteps, every, recompute):
        segs = _segments(int(nsteps), int(every))
        u0 = u0.contiguous()
        out = torch.empty((len(segs), u0.numel()), dtype=u0.dtype, device=u0.device)
        t0 = mesh.time
        _handoff(mesh)
        mesh.copy_in(_ptr(u0, mesh), REG_U)
        mesh.checkpoint_alloc(len(segs) + 1)
        for s, k in enumerate(segs):
            mesh.checkpoint_store(s, REG_U)
            check(lib().nmx_evolve(mesh.ctx, k))
            mesh.copy_out(_ptr(out[s], mesh), REG_U)
        mesh.checkpoint_store(len(segs), REG_U)   # the final state, restored after backward
        mesh.synchronize()
        ctx.mesh = mesh
        ctx.segs = segs
        ctx.recompute = bool(recompute)
        ctx.t0 = t0
        return out

    @staticmethod
    def backward(ctx, grad_out):
        mesh, segs = ctx.mesh, ctx.segs
        grad_out = grad_out.contiguous()
        lam = grad_out[-1].clone()
        _handoff(mesh)
        for s in range(len(segs) - 1, -1, -1):
            if ctx.recompute:
                # the segment's forward states, from its checkpoint
                mesh.checkpoint_load(s, REG_U)
                check(lib().nmx_evolve(mesh.ctx, segs[s]))
            mesh.copy_in(_ptr(lam, mesh), REG_ADJ)
            check(lib().nmx_adjoint_evolve(mesh.ctx, segs[s]))
            mesh.copy_out(_ptr(lam, mesh), REG_ADJ)
            mesh.synchronize()
            if s > 0:
                lam += grad_out[s - 1]
                torch.cuda.current_stream().synchronize()
        if ctx.recompute:
            mesh.checkpoint_load(len(segs), REG_U)
            mesh.synchronize()
        return lam, None, None, None, None


def evolve(mesh, u0, nsteps, every=16, recompute=False):
    """Differentiable evolution of u0 (flat float64 GPU tensor); returns the
    states after each segment of ``every`` steps, the last row being
    u(nsteps)."""
    return EvolveFunction.apply(u0, mesh, int(nsteps), int(every), bool(recompute))
