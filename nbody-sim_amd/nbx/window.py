"""Differentiable window of the collision-free disk step.

The forward pass is the library's own step loop (main.cu:85-140 with
collisions off): KDK, gas drag, monopole Barnes-Hut walk, star and
perturber.  The backward pass is the hand-written reverse mode of the same
float step (nbx_adj_step / nbx_adj_init_forces): FP32 arithmetic per
interaction, FP64 adjoint state and sums, and the forward walk's interaction
lists (the adjoint walk takes the forward walk's decisions on bit-identical
positions).  Memory follows checkpoint-every-k recomputation: the forward
keeps the state every k steps, the backward re-runs each segment of k steps
once, keeping its k + 1 states, and reverses it.

DiskWindow.apply(x0, v0, pert_mass_mjup, pert_radius_au, mass, tree, cfg)
returns the final (x, v), each (3, N) float32 on the device, and
differentiates with respect to x0, v0, the perturber mass (Jupiter masses)
and orbit radius (AU).
"""
import ctypes as C
from dataclasses import dataclass

import torch

from . import (MJUP_TO_MSUN, ParticleData, _check, _disk_array, disk_params, init_forces, lib,
               makePerturber, set_precision, state64, step, step_params)


@dataclass
class WindowConfig:
    n_steps: int = 200
    checkpoint_every: int = 20
    dt: float = 0.01
    theta: float = 0.5
    softening: float = 1e-3
    star_mass: float = 1.0
    gas_drag_rate: float = 1e-3
    pert_phase0: float = 0.0


def _state(x, v, a, mass, ids):
    """ParticleData view of (3, N) position / velocity / acceleration tensors."""
    p = ParticleData(x.shape[1])
    p.t = {"x": x[0], "y": x[1], "z": x[2], "vx": v[0], "vy": v[1], "vz": v[2],
           "ax": a[0], "ay": a[1], "az": a[2], "mass": mass, "id": ids}
    return p


def _params(cfg, pert_mass_mjup, pert_radius_au):
    sp = step_params(dt=cfg.dt, theta=cfg.theta, softening=cfg.softening,
                     star_mass=cfg.star_mass, restitution=0.0, gas_drag_rate=cfg.gas_drag_rate,
                     collisions=False, use_graph=False)
    pert = makePerturber(True, float(pert_mass_mjup), float(pert_radius_au), cfg.pert_phase0,
                         cfg.star_mass)
    return sp, disk_params(0.0, pert)


def _check_inputs(x0, v0, mass, tree):
    if x0.dim() != 2 or x0.shape[0] != 3 or v0.shape != x0.shape:
        raise RuntimeError("x0 and v0 must be (3, N) tensors")
    if x0.dtype != torch.float32 or v0.dtype != torch.float32 or mass.dtype != torch.float32:
        raise RuntimeError("window state is float32")
    if not (x0.is_cuda and v0.is_cuda and mass.is_cuda):
        raise RuntimeError("window state must be on the HIP device")
    if mass.shape != (x0.shape[1],):
        raise RuntimeError("mass must be (N,)")
    if tree.n_bodies != x0.shape[1]:
        raise RuntimeError(f"tree holds {tree.n_bodies} bodies, state has {x0.shape[1]}")


def run_window(x0, v0, pert_mass_mjup, pert_radius_au, mass, tree, cfg, keep=None):
    """Forward window without gradients: returns (x, v, a) after cfg.n_steps.
    keep (a list) receives the (x, v, a) checkpoints at steps 0, k, 2k, ..."""
    _check_inputs(x0, v0, mass, tree)
    sp, dp = _params(cfg, pert_mass_mjup, pert_radius_au)
    n = x0.shape[1]
    ids = torch.arange(n, dtype=torch.int32, device=x0.device)
    x = x0.detach().clone().contiguous()
    v = v0.detach().clone().contiguous()
    a = torch.zeros_like(x)
    S = _state(x, v, a, mass, ids)
    init_forces(tree, S, sp, dp)
    k = cfg.checkpoint_every
    for s0 in range(0, cfg.n_steps, k):
        if keep is not None:
            keep.append((x.clone(), v.clone(), a.clone()))
        step(tree, S, sp, dp, s0, min(k, cfg.n_steps - s0))
    return x, v, a


def run_window64(x0, v0, pert_mass_mjup, pert_radius_au, mass, tree, cfg):
    """The same forward window through the all-FP64 step (FP64 positions,
    velocities, moments and forces; the octree from the rounded positions):
    returns the final FP64 state (x, v), each (3, N) float64.  Its rounding
    (~1e-16 relative) leaves central differences with a relative step of 1e-4
    clean, where the float window's rounding (~1e-7 per step) does not.  The
    tree is switched to the FP64 step for the run and back to the float step
    afterwards."""
    _check_inputs(x0, v0, mass, tree)
    sp, dp = _params(cfg, pert_mass_mjup, pert_radius_au)
    n = x0.shape[1]
    ids = torch.arange(n, dtype=torch.int32, device=x0.device)
    x = x0.detach().clone().contiguous()
    v = v0.detach().clone().contiguous()
    S = _state(x, v, torch.zeros_like(x), mass, ids)
    set_precision(tree, "fp64")
    try:
        init_forces(tree, S, sp, dp)  # loads the FP64 state from x0, v0
        step(tree, S, sp, dp, 0, cfg.n_steps)
        s = state64(tree, S)
    finally:
        set_precision(tree, "float")
    return s[0:3].clone(), s[3:6].clone()


class DiskWindow(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x0, v0, pert_mass_mjup, pert_radius_au, mass, tree, cfg):
        if cfg.checkpoint_every < 1 or cfg.n_steps < 1:
            raise RuntimeError("n_steps and checkpoint_every must be positive")
        _check(lib().nbx_tree_set_adjoint(tree.h, 1), "set_adjoint")
        ckpts = []
        x, v, _ = run_window(x0, v0, float(pert_mass_mjup), float(pert_radius_au), mass, tree,
                             cfg, keep=ckpts)
        ctx.ckpts = ckpts
        ctx.tree, ctx.cfg, ctx.mass = tree, cfg, mass
        ctx.pert = (float(pert_mass_mjup), float(pert_radius_au))
        ctx.pdtype = (pert_mass_mjup.dtype if torch.is_tensor(pert_mass_mjup) else torch.float64,
                      pert_radius_au.dtype if torch.is_tensor(pert_radius_au) else torch.float64)
        return x, v

    @staticmethod
    def backward(ctx, gx, gv):
        tree, cfg, mass = ctx.tree, ctx.cfg, ctx.mass
        sp, dp = _params(cfg, *ctx.pert)
        dpa = _disk_array(dp, tree)
        x_ref = ctx.ckpts[0][0]
        n = x_ref.shape[1]
        ids = torch.arange(n, dtype=torch.int32, device=x_ref.device)
        zeros = torch.zeros(3, n, dtype=torch.float64, device=x_ref.device)
        xb = gx.double().contiguous() if gx is not None else zeros.clone()
        vb = gv.double().contiguous() if gv is not None else zeros.clone()
        ab = zeros.clone()
        pg = (C.c_double * 2)(0.0, 0.0)
        L = lib()
        k = cfg.checkpoint_every
        for c in reversed(range(len(ctx.ckpts))):
            s0 = c * k
            m = min(k, cfg.n_steps - s0)
            x, v, a = (t.clone() for t in ctx.ckpts[c])
            S = _state(x, v, a, mass, ids)
            states = [(x.clone(), v.clone(), a.clone())]
            for j in range(m):  # recompute the segment, bit-identical to the forward
                step(tree, S, sp, dp, s0 + j, 1)
                states.append((x.clone(), v.clone(), a.clone()))
            for j in reversed(range(m)):
                S0 = _state(*states[j], mass, ids)
                S1 = _state(*states[j + 1], mass, ids)
                _check(L.nbx_adj_step(tree.h, C.byref(S0.c()), C.byref(S1.c()), C.byref(sp),
                                      dpa, s0 + j, C.c_void_p(xb.data_ptr()),
                                      C.c_void_p(vb.data_ptr()), C.c_void_p(ab.data_ptr()), pg),
                       "adj_step")
        S0 = _state(*ctx.ckpts[0], mass, ids)
        _check(L.nbx_adj_init_forces(tree.h, C.byref(S0.c()), C.byref(sp), dpa,
                                     C.c_void_p(ab.data_ptr()), C.c_void_p(xb.data_ptr()), pg),
               "adj_init_forces")
        dm = torch.tensor(pg[0] * MJUP_TO_MSUN, dtype=ctx.pdtype[0])
        dr = torch.tensor(pg[1], dtype=ctx.pdtype[1])
        return xb.float(), vb.float(), dm, dr, None, None, None


def window(x0, v0, pert_mass_mjup, pert_radius_au, mass, tree, cfg):
    """(x, v) after cfg.n_steps collision-free steps, differentiable in x0,
    v0, the perturber mass (M_Jup) and radius (AU)."""
    return DiskWindow.apply(x0, v0, pert_mass_mjup, pert_radius_au, mass, tree, cfg)


def orbital_elements(x, v, mu=1.0):
    """Squared eccentricity and inclination (rad) of heliocentric orbits.
    x, v: (3, N).  e = ((v^2 - mu/r) r - (r.v) v) / mu; prograde orbits of
    this disk (XZ plane, DiskInit.cpp:50-61) have h along -Y, so the
    inclination is the angle between h and -Y."""
    r = torch.sqrt((x * x).sum(0))
    v2 = (v * v).sum(0)
    rv = (x * v).sum(0)
    ev = ((v2 - mu / r) * x - rv * v) / mu
    e2 = (ev * ev).sum(0)
    h = torch.linalg.cross(x, v, dim=0)
    inc = torch.atan2(torch.sqrt(h[0] * h[0] + h[2] * h[2]), -h[1])
    return e2, inc


def annulus_mask(x, r_lo=1.0, r_hi=2.0):
    r = torch.sqrt((x.double() ** 2).sum(0))
    return (r >= r_lo) & (r <= r_hi)


def window_loss(x, v, mask, mu=1.0):
    """Mean of e^2 + i^2 over the bodies in mask (evaluated in FP64)."""
    e2, inc = orbital_elements(x.double(), v.double(), mu)
    return (e2[mask] + inc[mask] ** 2).mean()
