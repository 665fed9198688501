"""torch.autograd.Function over differentiable stepping (cc_fwd / cc_bwd through the C ABI).

    com = WorldRun.apply(v0, friction, context, scene, dt, steps, checkpoint_every)

runs `steps` steps of dt on `context` from `scene` with its initial linear velocities replaced
by v0 ((n, 3)) and every body's friction coefficient by friction ((n,)), and returns the
centres of mass after the run ((n, 3), float64 on the CPU).  Its backward takes dL/dcom and
returns dL/dv0 and dL/dfriction from the adjoint of the velocity-level contact solve and the
integrator (include/crashcat_amd.h, cc_bwd).  Tensors cross the boundary as host arrays; the
HIP library does all the stepping.
"""
import numpy as np
import torch


class WorldRun(torch.autograd.Function):
    @staticmethod
    def forward(ctx, v0, friction, context, scene, dt, steps, checkpoint_every=10):
        b = scene.bodies()
        keep = b.copy()
        try:
            b["linear_velocity"][:] = v0.detach().cpu().numpy().astype(np.float64).reshape(-1, 3)
            b["friction"][:] = friction.detach().cpu().numpy().astype(np.float64).reshape(-1)
            context.upload(scene)
        finally:
            b[:] = keep
        context.fwd(dt, steps, checkpoint_every)
        com = context.bodies()["center_of_mass_position"].copy()
        ctx.context = context
        return torch.from_numpy(com)

    @staticmethod
    def backward(ctx, grad_com):
        out = ctx.context.bwd(grad_com.detach().cpu().numpy().astype(np.float64))
        return (torch.from_numpy(out["v0"]), torch.from_numpy(out["friction"]), None, None, None, None, None)
