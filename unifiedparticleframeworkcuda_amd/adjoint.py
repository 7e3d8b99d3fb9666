"""Differentiable PCISPH rollout (upf_adjoint_* in include/upf.h) and its
torch.autograd.Function wrapper for gradient checks.

The rollout runs on the HIP device through the C ABI; torch only carries
host (CPU) tensors in and out of it -- the library keeps its own device
buffers and stream.  The step is the static-boundary PCISPH step of the CPU
path (src/cpu/UnifiedPhysics.cpp:4220-4238) with a fixed iteration count and
closed-form kernels (see the header); the backward pass is its exact
adjoint, with the state kept every `checkpoint_every` steps and each segment
recomputed during the backward pass.
"""
import ctypes as C

import numpy as np

from . import check, lib, standard_dam_break, UPF_LIQUID
from ._abi import Params, ptr


def _bind():
    L = lib()
    if getattr(L, "_adjoint_bound", False):
        return L
    vp = C.c_void_p
    pf = C.POINTER(C.c_float)
    L.upf_adjoint_create.argtypes = [C.POINTER(Params), C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(vp)]
    L.upf_adjoint_destroy.argtypes = [vp]
    L.upf_adjoint_forward.argtypes = [vp, pf, pf, C.POINTER(C.c_int), pf, pf]
    L.upf_adjoint_backward.argtypes = [vp, pf, pf, pf, pf]
    L.upf_adjoint_timing.argtypes = [vp, pf, pf, C.POINTER(C.c_size_t)]
    L._adjoint_bound = True
    return L


def _f4(a):
    a = np.asarray(a, np.float32)
    out = np.zeros((len(a), 4), np.float32)
    out[:, :3] = a[:, :3]
    return out


class Rollout:
    """`steps` differentiable PCISPH steps with `iterations` correction
    iterations each, for a fixed particle set (positions, types)."""

    def __init__(self, params, pos, types, steps, iterations=5, checkpoint_every=8):
        self.L = _bind()
        self.p = params.copy()
        self.pos = _f4(pos)
        self.types = np.ascontiguousarray(types, np.int32)
        self.n = len(self.types)
        self.steps = steps
        self.h = C.c_void_p()
        check("upf_adjoint_create", self.L.upf_adjoint_create(C.byref(self.p), self.n, int(steps), int(iterations),
                                                              int(checkpoint_every), C.byref(self.h)))

    def forward(self, vel, pos=None):
        """(x_final, v_final) as (n, 3) float32 arrays."""
        x0 = self.pos if pos is None else _f4(pos)
        v0 = _f4(vel)
        xo = np.zeros((self.n, 4), np.float32)
        vo = np.zeros((self.n, 4), np.float32)
        check("upf_adjoint_forward", self.L.upf_adjoint_forward(self.h, ptr(x0), ptr(v0), ptr(self.types, C.c_int),
                                                                ptr(xo), ptr(vo)))
        return xo[:, :3].copy(), vo[:, :3].copy()

    def backward(self, dl_dx, dl_dv=None):
        """(dL/dx0, dL/dv0) for the last forward call."""
        gx = _f4(dl_dx)
        gv = None if dl_dv is None else _f4(dl_dv)
        ox = np.zeros((self.n, 4), np.float32)
        ov = np.zeros((self.n, 4), np.float32)
        check("upf_adjoint_backward", self.L.upf_adjoint_backward(self.h, ptr(gx), ptr(gv), ptr(ox), ptr(ov)))
        return ox[:, :3].copy(), ov[:, :3].copy()

    def timing(self):
        """(forward ms, backward ms, device bytes) of the last calls."""
        f, b, m = C.c_float(), C.c_float(), C.c_size_t()
        check("upf_adjoint_timing", self.L.upf_adjoint_timing(self.h, C.byref(f), C.byref(b), C.byref(m)))
        return f.value, b.value, m.value

    def close(self):
        if self.h:
            self.L.upf_adjoint_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def differentiable_dam_break(nx=64, ny=32, nz=32, v_sigma=0.1, target_sigma=0.02, seed_v=4, seed_target=5,
                             spacing=0.01, dt=0.001):
    """The dam break of the gradient workload: an nx*ny*nz lattice (config-0
    layout), initial fluid velocities N(0, v_sigma) per component (seed_v),
    and a target point cloud of the same size: the initial fluid positions
    plus N(0, target_sigma) (seed_target).  Returns (params, pos4, types,
    vel (n, 3), target (n, 3), fluid mask)."""
    p, pos, typ, nf = standard_dam_break(nx, ny, nz, spacing=spacing, dt=dt)
    f = typ == UPF_LIQUID
    vel = np.zeros((len(typ), 3), np.float32)
    vel[f] = np.random.RandomState(seed_v).normal(0.0, v_sigma, size=(int(f.sum()), 3))
    target = pos[:, :3].astype(np.float32).copy()
    target[f] += np.random.RandomState(seed_target).normal(0.0, target_sigma, size=(int(f.sum()), 3))
    return p, pos, typ, vel, target, f


def torch_rollout_function():
    """torch.autograd.Function: v0 (n, 3) CPU tensor -> x_final (n, 3)."""
    import torch

    class PCISPHRollout(torch.autograd.Function):
        @staticmethod
        def forward(ctx, v0, rollout):
            x, _ = rollout.forward(v0.detach().cpu().numpy())
            ctx.rollout = rollout
            return torch.from_numpy(x)

        @staticmethod
        def backward(ctx, grad_x):
            _, gv = ctx.rollout.backward(grad_x.detach().cpu().numpy())
            return torch.from_numpy(gv), None

    return PCISPHRollout
