"""CPU checks of the differentiable-window pieces that need no GPU: the
orbital-element loss and the FP64 PyTorch oracle the GPU reverse mode is
checked against (pinned here by closed forms, central differences and the
C restatement of the reference step loop)."""
import math

import numpy as np
import pytest
import torch

import pyoracle as O
import window_ref as W
from helpers import disk
from nbx.window import WindowConfig, orbital_elements, window_loss

F64 = torch.float64


def kepler_state(a, e, inc, argp, node, nu, mu=1.0):
    """Position / velocity of a Keplerian orbit whose reference plane is XZ
    with prograde angular momentum along -Y (the disk's convention)."""
    p = a * (1 - e * e)
    r = p / (1 + e * math.cos(nu))
    # perifocal frame: P toward pericentre, Q 90 deg ahead, W = angular momentum
    rp = np.array([r * math.cos(nu), r * math.sin(nu), 0.0])
    vp = np.sqrt(mu / p) * np.array([-math.sin(nu), e + math.cos(nu), 0.0])

    def rot(axis, ang):
        c, s = math.cos(ang), math.sin(ang)
        if axis == 3:
            return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])
        return np.array([[1, 0, 0], [0, c, -s], [0, s, c]])
    R = rot(3, node) @ rot(1, inc) @ rot(3, argp)
    # map the textbook frame (x, y, z=h) to the disk frame (x, z, -y)
    M = np.array([[1, 0, 0], [0, 0, -1], [0, 1, 0]])
    return M @ R @ rp, M @ R @ vp


@pytest.mark.parametrize("e,inc", [(0.0, 0.0), (0.1, 0.0), (0.3, 0.2), (0.05, 1.0), (0.6, 2.5)])
def test_orbital_elements_closed_form(e, inc):
    for nu in (0.0, 1.0, 2.5, 4.0):
        x, v = kepler_state(1.7, e, inc, 0.4, 1.3, nu)
        e2, i = orbital_elements(torch.tensor(x)[:, None], torch.tensor(v)[:, None])
        assert abs(e2.item() - e * e) < 1e-12
        assert abs(i.item() - inc) < 1e-12


def test_disk_orbits_are_prograde_about_minus_y():
    d = disk(2000, seed=3)
    x = torch.tensor(np.stack([d["x"], d["y"], d["z"]]), dtype=torch.float64)
    v = torch.tensor(np.stack([d["vx"], d["vy"], d["vz"]]), dtype=torch.float64)
    e2, i = orbital_elements(x, v)
    # generator: Rayleigh e (sigma 0.01) and i (sigma 0.005), vertical thickness 0.01 r
    assert np.median(np.sqrt(e2.numpy())) < 0.05
    assert np.median(i.numpy()) < 0.05


def small_case(n=24, seed=5):
    d = disk(n, mass=3e-3, seed=seed)  # heavy bodies: self-gravity matters
    x = torch.tensor(np.stack([d["x"], d["y"], d["z"]]).astype(np.float64))
    v = torch.tensor(np.stack([d["vx"], d["vy"], d["vz"]]).astype(np.float64))
    m = torch.tensor(d["mass"].astype(np.float64))
    return x, v, m


def loss_of(cfg, x0, v0, m, pm, rp, **kw):
    x, v = W.window(x0, v0, m, pm, rp, cfg, **kw)
    mask = torch.ones(x.shape[1], dtype=torch.bool)
    return window_loss(x, v, mask)


def test_reference_autograd_matches_central_differences():
    cfg = WindowConfig(n_steps=12, dt=0.01, softening=1e-3, gas_drag_rate=1e-2)
    x0, v0, m = small_case()
    pm = torch.tensor(10 * W.MJUP_TO_MSUN, dtype=F64, requires_grad=True)
    rp = torch.tensor(5.2, dtype=F64, requires_grad=True)
    v0g = v0.clone().requires_grad_(True)
    x0g = x0.clone().requires_grad_(True)
    L = loss_of(cfg, x0g, v0g, m, pm, rp)
    L.backward()

    def fd(f, h):
        return (f(h) - f(-h)) / (2 * h)
    h_pm = 1e-4 * pm.item()
    g_pm = fd(lambda h: loss_of(cfg, x0, v0, m, torch.tensor(pm.item() + h, dtype=F64), rp.detach()).item(),
              h_pm)
    h_rp = 1e-4 * rp.item()
    g_rp = fd(lambda h: loss_of(cfg, x0, v0, m, pm.detach(), torch.tensor(rp.item() + h, dtype=F64)).item(),
              h_rp)
    assert abs(pm.grad.item() - g_pm) <= 1e-5 * abs(g_pm)
    assert abs(rp.grad.item() - g_rp) <= 1e-5 * abs(g_rp)
    for (c, j) in [(0, 3), (2, 7), (1, 11)]:
        def fv(h):
            vv = v0.clone()
            vv[c, j] += h
            return loss_of(cfg, x0, vv, m, pm.detach(), rp.detach()).item()
        def fx(h):
            xx = x0.clone()
            xx[c, j] += h
            return loss_of(cfg, xx, v0, m, pm.detach(), rp.detach()).item()
        gv = fd(fv, 1e-4 * abs(v0[c, j].item()) + 1e-8)
        gx = fd(fx, 1e-4 * abs(x0[c, j].item()) + 1e-8)
        assert abs(v0g.grad[c, j].item() - gv) <= 1e-5 * abs(gv) + 1e-9
        assert abs(x0g.grad[c, j].item() - gx) <= 1e-5 * abs(gx) + 1e-9


def test_reference_window_follows_reference_step_loop():
    """FP64 direct-sum window == the C restatement of main.cu:85-140 with a
    vanishing opening angle (every node opened: per-body direct sum)."""
    n, steps = 64, 10
    d = disk(n, mass=1e-5, seed=11)
    cfg = WindowConfig(n_steps=steps, dt=0.01, softening=1e-3, gas_drag_rate=1e-3)
    params = O.step_params(dt=0.01, theta=1e-6, softening=1e-3, star_mass=1.0, r_coll=0.0,
                           gas_drag_rate=1e-3, perturber_enabled=1, pert_mass_mjup=10.0,
                           pert_radius_au=5.2, collisions=0)
    s = O.ref_steps(d, d["mass"], params, 0, steps, init_forces=True, nthreads=1)
    x0 = torch.tensor(np.stack([d["x"], d["y"], d["z"]]).astype(np.float64))
    v0 = torch.tensor(np.stack([d["vx"], d["vy"], d["vz"]]).astype(np.float64))
    m = torch.tensor(d["mass"].astype(np.float64))
    pm = torch.tensor(float(np.float32(np.float32(10.0) * np.float32(W.MJUP_TO_MSUN))), dtype=F64)
    x, v = W.window(x0, v0, m, pm, torch.tensor(5.2, dtype=F64), cfg, checkpoint=False)
    xr = np.stack([s["x"], s["y"], s["z"]])
    vr = np.stack([s["vx"], s["vy"], s["vz"]])
    assert np.max(np.abs(x.numpy() - xr)) < 2e-6
    assert np.max(np.abs(v.numpy() - vr)) < 2e-6


def test_list_force_equals_direct_sum_when_every_node_opens():
    x, _, m = small_case(n=40, seed=2)
    xf = x.numpy().astype(np.float32)
    perm, off, lo, hi = O.mt_lists(xf[0], xf[1], xf[2], m.numpy().astype(np.float32), 1e-7)
    assert np.all(lo == hi)  # leaves only
    a1 = W.accel_lists(x, m, torch.tensor(perm).long(), torch.tensor(off), torch.tensor(lo).long(),
                       torch.tensor(hi).long(), 1e-3)
    a2 = W.accel_direct(x, m, 1e-3)
    assert torch.allclose(a1, a2, rtol=1e-12, atol=1e-15)


def test_list_force_matches_oracle_walk():
    d = disk(3000, mass=1e-5, seed=4)
    f32 = [d[k] for k in ("x", "y", "z", "mass")]
    perm, off, lo, hi = O.mt_lists(*f32, 0.5)
    ref = O.mt_force(*f32, 0.5, 1e-3)["self_a"]
    x = torch.tensor(np.stack(f32[:3]).astype(np.float64))
    m = torch.tensor(d["mass"].astype(np.float64))
    a = W.accel_lists(x, m, torch.tensor(perm).long(), torch.tensor(off), torch.tensor(lo).long(),
                      torch.tensor(hi).long(), 1e-3).numpy().T
    err = np.linalg.norm(a - ref, axis=1) / np.linalg.norm(ref, axis=1)
    assert np.median(err) < 1e-5 and np.max(err) < 1e-3


def test_list_force_gradient_matches_central_differences():
    """The node-COM chain (X_c = sum m_j x_j / M_c) of the frozen-list force."""
    x, _, m = small_case(n=60, seed=8)
    xf = x.numpy().astype(np.float32)
    perm, off, lo, hi = (torch.tensor(a) for a in O.mt_lists(*xf, m.numpy().astype(np.float32),
                                                             0.7))
    perm, lo, hi = perm.long(), lo.long(), hi.long()
    assert torch.any(lo != hi)  # some internal nodes are accepted
    w = torch.randn(3, x.shape[1], generator=torch.Generator().manual_seed(0), dtype=F64)
    f = lambda y: (W.accel_lists(y, m, perm, off, lo, hi, 1e-3) * w).sum()
    xg = x.clone().requires_grad_(True)
    f(xg).backward()
    for (c, j) in [(0, 1), (1, 30), (2, 59), (0, 17)]:
        h = 1e-6
        xp, xm = x.clone(), x.clone()
        xp[c, j] += h
        xm[c, j] -= h
        g = (f(xp) - f(xm)).item() / (2 * h)
        assert abs(xg.grad[c, j].item() - g) <= 1e-6 * abs(g) + 1e-10
