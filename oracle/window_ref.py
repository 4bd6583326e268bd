"""FP64 PyTorch restatement of the collision-free disk window (test oracle).

Test infrastructure only: the tests import it to check the library's
hand-written reverse mode; no product path uses it.

The step order is main.cu:85-140 with collisions off:
  initial forces at t = 0 (main.cu:85-96), then per step n
  v += a dt/2; x += v dt (Integrator.cu:7-23); a = A(x, t = (n+1) dt)
  (BHTree.cu:530-609 star + self-gravity, Perturber.cu:131-143);
  v += a dt/2; gas drag (Integrator.cu:40-76).
Self-gravity is either the exact softened direct sum or a frozen list of
accepted tree nodes whose centres of mass are mass-weighted means of the
member positions (so autograd differentiates through them).  Everything is
float64 and differentiable, including the perturber orbit
r_p (cos phi, 0, sin phi), phi = phi0 + sqrt(M / r_p^3) t (Perturber.h:39-66).
"""
import math

import torch
import torch.utils.checkpoint as ckpt

MJUP_TO_MSUN = 9.5479e-4  # Perturber.h:25


def pull(d, s):
    """Softened pull d / s^{3/2} (d: (3, ...), s = |d|^2 + eps^2)."""
    return d * s.pow(-1.5)


def accel_direct(x, m, eps):
    """Exact softened self-gravity: a_i = sum_{j != i} m_j (x_j - x_i) / (r^2 + eps^2)^{3/2}."""
    n = x.shape[1]
    d = x[:, None, :] - x[:, :, None]  # d[:, i, j] = x_j - x_i
    s = (d * d).sum(0) + eps * eps
    w = m[None, :] * s.pow(-1.5) * (1.0 - torch.eye(n, dtype=x.dtype, device=x.device))
    return (d * w[None]).sum(2)


def accel_lists(x, m, perm, off, lo, hi, eps):
    """Self-gravity over frozen interaction lists: target i pulls towards the
    COM of each accepted node, the node being the sorted positions
    [lo, hi] of perm (orc_mt_lists)."""
    n = x.shape[1]
    xs = x[:, perm]
    ms = m[perm]
    pm = torch.cat([torch.zeros(1, dtype=x.dtype, device=x.device), torch.cumsum(ms, 0)])
    pmx = torch.cat([torch.zeros(3, 1, dtype=x.dtype, device=x.device),
                     torch.cumsum(ms[None, :] * xs, 1)], 1)
    mc = pm[hi + 1] - pm[lo]
    xc = (pmx[:, hi + 1] - pmx[:, lo]) / mc[None, :]
    tgt = torch.repeat_interleave(torch.arange(n, device=x.device), off[1:] - off[:-1])
    d = xc - x[:, tgt]
    s = (d * d).sum(0) + eps * eps
    f = d * (mc * s.pow(-1.5))[None, :]
    a = torch.zeros_like(x)
    return a.index_add(1, tgt, f)


def perturber_pos(r_p, phase0, star_mass, t):
    omega = torch.sqrt(star_mass / (r_p * r_p * r_p))
    phi = phase0 + omega * t
    z = torch.zeros_like(r_p)
    return torch.stack([r_p * torch.cos(phi), z, r_p * torch.sin(phi)])


def accel(x, m, pmass, r_p, t, cfg, self_grav):
    eps = cfg.softening
    a = self_grav(x)
    s = (x * x).sum(0) + eps * eps
    a = a - cfg.star_mass * pull(x, s)
    xp = perturber_pos(r_p, cfg.pert_phase0, cfg.star_mass, t)
    d = xp[:, None] - x
    a = a + pmass * pull(d, (d * d).sum(0) + eps * eps)
    return a


def drag(x, v, decay, star_mass):
    """Integrator.cu:40-76: radial and vertical velocity damped by `decay`,
    azimuthal velocity relaxed towards the signed Keplerian speed."""
    r = torch.sqrt(x[0] * x[0] + x[2] * x[2])
    ux, uz = x[0] / r, x[2] / r
    vr = v[0] * ux + v[2] * uz
    vphi = -v[0] * uz + v[2] * ux
    vk = torch.sqrt(star_mass / r) * torch.where(vphi >= 0, 1.0, -1.0).to(x.dtype)
    vr2 = vr * decay
    vphi2 = vk + (vphi - vk) * decay
    return torch.stack([vr2 * ux - vphi2 * uz, v[1] * decay, vr2 * uz + vphi2 * ux])


def window(x0, v0, m, pmass_msun, r_p, cfg, self_grav="direct", lists=None, checkpoint=True):
    """(x, v) after cfg.n_steps steps, all float64 and differentiable.
    self_grav: "direct", or "lists" with lists(x) -> (perm, off, lo, hi) of
    the walk on those positions (frozen: the lists are data, not variables)."""
    eps = cfg.softening
    dt, dth = cfg.dt, 0.5 * cfg.dt
    decay = math.exp(-cfg.gas_drag_rate * cfg.dt) if cfg.gas_drag_rate > 0 else 1.0

    def sg_for(x):
        if self_grav == "direct":
            return lambda y: accel_direct(y, m, eps)
        perm, off, lo, hi = lists(x.detach())
        return lambda y: accel_lists(y, m, perm, off, lo, hi, eps)

    def force(x, t):
        sg = sg_for(x)
        f = lambda y, pm, rp: accel(y, m, pm, rp, t, cfg, sg)
        if checkpoint:
            return ckpt.checkpoint(f, x, pmass_msun, r_p, use_reentrant=False)
        return f(x, pmass_msun, r_p)

    x, v = x0, v0
    a = force(x, 0.0)
    for n in range(cfg.n_steps):
        v = v + a * dth
        x = x + v * dt
        a = force(x, (n + 1) * dt)
        v = v + a * dth
        if cfg.gas_drag_rate > 0:
            v = drag(x, v, decay, cfg.star_mass)
    return x, v
