"""Reverse mode of the collision-free step on the GPU (nbx_adj_force,
nbx_adj_step, nbx_adj_init_forces through DiskWindow) against the FP64
PyTorch oracle (frozen interaction lists of the oracle walk, or the exact
direct sum when every node opens) and against central differences of the
library's own forward window."""
import ctypes as C

import numpy as np
import pytest
import torch

import pyoracle as O
import window_ref as W
from helpers import disk

pytestmark = pytest.mark.gpu

F64 = torch.float64


def dev_state(nbx, d):
    x = torch.tensor(np.stack([d["x"], d["y"], d["z"]]), device="cuda").contiguous()
    v = torch.tensor(np.stack([d["vx"], d["vy"], d["vz"]]), device="cuda").contiguous()
    m = torch.tensor(d["mass"], device="cuda")
    return x, v, m


def particles(nbx, x, v, a, m):
    from nbx.window import _state
    ids = torch.arange(x.shape[1], dtype=torch.int32, device="cuda")
    return _state(x, v, a, m, ids)


def adj_force(nbx, t, p, theta, dp, sim_time, ab):
    xb = torch.zeros_like(ab)
    pg = (C.c_double * 2)(0.0, 0.0)
    rc = nbx.lib().nbx_adj_force(t.h, C.byref(p.c()), theta, 1e-3, 1.0, C.byref(dp), sim_time,
                                 C.c_void_p(ab.data_ptr()), C.c_void_p(xb.data_ptr()), pg)
    assert rc == 0, nbx.lib().nbx_last_error()
    torch.cuda.synchronize()
    return xb, (pg[0], pg[1])


def torch_force_vjp(x, m, ab, theta, dp, sim_time, lists):
    """FP64 autograd of sum(ab * A(x)) on the CPU: A = self-gravity over the
    oracle walk's frozen lists (or direct sum) + star + perturber."""
    from nbx.window import WindowConfig
    cfg = WindowConfig(softening=1e-3, pert_phase0=dp.pert_phase0)
    xg = x.clone().requires_grad_(True)
    pm = torch.tensor(dp.pert_mass_msun, dtype=F64, requires_grad=True)
    rp = torch.tensor(dp.pert_radius_au, dtype=F64, requires_grad=True)
    if lists is None:
        sg = lambda y: W.accel_direct(y, m, 1e-3)
    else:
        perm, off, lo, hi = lists
        sg = lambda y: W.accel_lists(y, m, perm, off, lo, hi, 1e-3)
    a = W.accel(xg, m, pm, rp, float(np.float32(sim_time)), cfg, sg)
    (a * ab).sum().backward()
    return xg.grad, (pm.grad.item(), rp.grad.item())


def oracle_lists(d, theta):
    perm, off, lo, hi = O.mt_lists(d["x"], d["y"], d["z"], d["mass"], theta)
    return (torch.tensor(perm).long(), torch.tensor(off), torch.tensor(lo).long(),
            torch.tensor(hi).long())


@pytest.mark.parametrize("theta", [0.5, 1e-6])
def test_force_adjoint_matches_fp64_autograd(gpu, theta):
    nbx = gpu
    n = 2048
    d = disk(n, mass=3e-5, seed=21)
    print(f"force adjoint n={n} theta={theta}", flush=True)
    x, v, m = dev_state(nbx, d)
    a = torch.zeros_like(x)
    p = particles(nbx, x, v, a, m)
    t = nbx.allocateOctree(n)
    nbx.lib().nbx_tree_set_adjoint(t.h, 1)
    pert = nbx.makePerturber(True, 10.0, 5.2, 0.3, 1.0)
    dp = nbx.disk_params(0.0, pert)
    g = torch.Generator().manual_seed(3)
    ab = torch.randn(3, n, generator=g, dtype=F64)
    xb, pg = adj_force(nbx, t, p, theta, dp, 0.37, ab.cuda())
    xc = torch.tensor(np.stack([d["x"], d["y"], d["z"]]).astype(np.float64))
    mc = torch.tensor(d["mass"].astype(np.float64))
    lists = None if theta < 1e-3 else oracle_lists(d, theta)
    ref, rpg = torch_force_vjp(xc, mc, ab, theta, dp, 0.37, lists)
    got = xb.cpu()
    err = torch.linalg.norm(got - ref) / torch.linalg.norm(ref)
    per = torch.linalg.norm(got - ref, dim=0) / torch.linalg.norm(ref, dim=0)
    print(f"  rel L2 {err:.3e}  median per-body {per.median():.3e}", flush=True)
    assert err < 1e-5
    assert per.median() < 1e-5
    assert abs(pg[0] - rpg[0]) <= 1e-5 * abs(rpg[0])
    assert abs(pg[1] - rpg[1]) <= 1e-5 * abs(rpg[1])


def test_self_gravity_adjoint_through_node_centres(gpu):
    """Heavy bodies, star and perturber switched off: only the tree terms
    (target side and node-COM chain) remain."""
    nbx = gpu
    n = 4096
    d = disk(n, mass=1.0, seed=22)
    print(f"self-gravity adjoint n={n}", flush=True)
    x, v, m = dev_state(nbx, d)
    p = particles(nbx, x, v, torch.zeros_like(x), m)
    t = nbx.allocateOctree(n)
    nbx.lib().nbx_tree_set_adjoint(t.h, 1)
    dp = nbx.disk_params(0.0, None)
    ab = torch.randn(3, n, generator=torch.Generator().manual_seed(4), dtype=F64)
    xb = torch.zeros_like(ab).cuda()
    pg = (C.c_double * 2)(0.0, 0.0)
    ab_dev = ab.cuda()
    rc = nbx.lib().nbx_adj_force(t.h, C.byref(p.c()), 0.5, 1e-3, 0.0, C.byref(dp), 0.0,
                                 C.c_void_p(ab_dev.data_ptr()), C.c_void_p(xb.data_ptr()), pg)
    assert rc == 0
    torch.cuda.synchronize()
    xc = torch.tensor(np.stack([d["x"], d["y"], d["z"]]).astype(np.float64)).requires_grad_(True)
    mc = torch.tensor(d["mass"].astype(np.float64))
    perm, off, lo, hi = oracle_lists(d, 0.5)
    assert torch.any(lo != hi)
    (W.accel_lists(xc, mc, perm, off, lo, hi, 1e-3) * ab).sum().backward()
    ref = xc.grad
    err = torch.linalg.norm(xb.cpu() - ref) / torch.linalg.norm(ref)
    print(f"  rel L2 {err:.3e}", flush=True)
    assert err < 1e-4


def window_case(nbx, n, seed, theta, steps, ckpt):
    from nbx.window import WindowConfig
    d = disk(n, mass=3e-5, seed=seed)
    x0, v0, m = dev_state(nbx, d)
    t = nbx.allocateOctree(n)
    cfg = WindowConfig(n_steps=steps, checkpoint_every=ckpt, theta=theta)
    return d, x0, v0, m, t, cfg


def test_window_gradient_matches_fp64_direct_sum_autograd(gpu):
    """theta -> 0 (every node opened): the GPU window is the FP32 direct-sum
    dynamics; its hand-written gradient must match FP64 autograd of the same
    dynamics in PyTorch (orbital-element loss of the 1-2 AU annulus)."""
    nbx = gpu
    from nbx.window import annulus_mask, window, window_loss
    n, steps = 2048, 40
    d, x0, v0, m, t, cfg = window_case(nbx, n, 31, 1e-6, steps, 10)
    print(f"window vs FP64 autograd n={n} steps={steps}", flush=True)
    pm = torch.tensor(10.0, dtype=F64, requires_grad=True)
    rp = torch.tensor(5.2, dtype=F64, requires_grad=True)
    v0g = v0.clone().requires_grad_(True)
    xT, vT = window(x0, v0g, pm, rp, m, t, cfg)
    mask = annulus_mask(xT)
    L = window_loss(xT, vT, mask)
    L.backward()
    # FP64 oracle on the GPU (direct sum, checkpointed forces)
    xr0 = x0.double()
    vr0 = v0.double().clone().requires_grad_(True)
    mr = m.double()
    pmr = torch.tensor(float(np.float32(np.float32(10.0) * np.float32(W.MJUP_TO_MSUN))),
                       dtype=F64, device="cuda", requires_grad=True)
    rpr = torch.tensor(5.2, dtype=F64, device="cuda", requires_grad=True)
    xr, vr = W.window(xr0, vr0, mr, pmr, rpr, cfg)
    Lr = window_loss(xr, vr, mask)
    Lr.backward()
    print(f"  loss {L.item():.6e} vs {Lr.item():.6e}", flush=True)
    assert abs(L.item() - Lr.item()) <= 1e-4 * abs(Lr.item())
    gv, gvr = v0g.grad.double(), vr0.grad
    err_v = (torch.linalg.norm(gv - gvr) / torch.linalg.norm(gvr)).item()
    g_m = pm.grad.item()
    g_mr = pmr.grad.item() * W.MJUP_TO_MSUN
    g_r, g_rr = rp.grad.item(), rpr.grad.item()
    print(f"  dL/dv0 rel L2 {err_v:.3e}; dL/dm {g_m:.6e} vs {g_mr:.6e}; "
          f"dL/dr {g_r:.6e} vs {g_rr:.6e}", flush=True)
    assert err_v < 1e-3
    assert abs(g_m - g_mr) <= 1e-3 * abs(g_mr)
    assert abs(g_r - g_rr) <= 1e-3 * abs(g_rr)


def test_window_gradient_matches_central_differences(gpu):
    """theta = 0.5: the adjoint of the library's own tree dynamics against
    central differences of its forward window (linear loss on the final
    state).  A difference quotient must stand above the float32 rounding of
    the state, which accumulates to ~1e-6 AU per body over 60 steps: the
    loss weights follow the perturber's pull (coherent over the bodies) and
    the steps are 20 % of the perturber mass and 2 % of its radius."""
    nbx = gpu
    from nbx.window import run_window, window
    n, steps = 4096, 60
    d, x0, v0, m, t, cfg = window_case(nbx, n, 41, 0.5, steps, 10)
    print(f"window vs central differences n={n} steps={steps}", flush=True)
    g = torch.Generator(device="cuda").manual_seed(9)
    # weights along the perturber's direction (x axis at phase 0): its pull
    # moves every body the same way, so the loss responds coherently while
    # the rounding of the state adds up incoherently
    wx = torch.zeros(3, n, device="cuda")
    wx[0] = 1.0
    wv = wx.clone()

    def loss(x, v):
        return (wx.double() * x.double()).sum() + (wv.double() * v.double()).sum()

    pm = torch.tensor(10.0, dtype=F64, requires_grad=True)
    rp = torch.tensor(5.2, dtype=F64, requires_grad=True)
    v0g = v0.clone().requires_grad_(True)
    xT, vT = window(x0, v0g, pm, rp, m, t, cfg)
    loss(xT, vT).backward()

    def fwd(pmv, rpv, vv):
        x, v, _ = run_window(x0, vv, pmv, rpv, m, t, cfg)
        return loss(x, v).item()

    # the final state is close to linear in the perturber mass, so a wide
    # step keeps the difference far above the float32 rounding of the state
    hm, hr = 0.2, 0.02
    fd_m = (fwd(10.0 * (1 + hm), 5.2, v0) - fwd(10.0 * (1 - hm), 5.2, v0)) / (2 * hm * 10.0)
    fd_r = (fwd(10.0, 5.2 * (1 + hr), v0) - fwd(10.0, 5.2 * (1 - hr), v0)) / (2 * hr * 5.2)
    dv = torch.randn(3, n, generator=g, device="cuda")
    eps = 1e-4
    fd_v = (fwd(10.0, 5.2, v0 + eps * dv) - fwd(10.0, 5.2, v0 - eps * dv)) / (2 * eps)
    ad_v = (v0g.grad.double() * dv.double()).sum().item()
    print(f"  dm {pm.grad.item():.6e} vs {fd_m:.6e}; dr {rp.grad.item():.6e} vs {fd_r:.6e}; "
          f"dv {ad_v:.6e} vs {fd_v:.6e}", flush=True)
    assert abs(pm.grad.item() - fd_m) <= 3e-2 * abs(fd_m)
    assert abs(rp.grad.item() - fd_r) <= 3e-2 * abs(fd_r)
    assert abs(ad_v - fd_v) <= 1e-2 * abs(fd_v)


def test_window_matches_fp64_central_differences_h1e4(gpu):
    """The orbital-element loss (mean e^2 + i^2 of the 1-2 AU annulus) at
    theta = 0.5: the hand-written gradient in the perturber mass, its radius
    and a random direction of v0 against central differences with a relative
    step of 1e-4, each forward run through the FP64 step (run_window64), whose
    rounding stays ~1e-16 where the float window's would swamp a 1e-4 step.
    The two forwards differ by float rounding (~1e-6 of the state), so their
    derivatives agree to well under 1 % (measured: 4e-4, 1e-4, 6e-4)."""
    nbx = gpu
    from nbx.window import annulus_mask, run_window64, window, window_loss
    n, steps = 4096, 60
    d, x0, v0, m, t, cfg = window_case(nbx, n, 43, 0.5, steps, 10)
    print(f"window vs FP64 central differences h=1e-4 n={n} steps={steps}", flush=True)
    pm = torch.tensor(10.0, dtype=F64, requires_grad=True)
    rp = torch.tensor(5.2, dtype=F64, requires_grad=True)
    v0g = v0.clone().requires_grad_(True)
    xT, vT = window(x0, v0g, pm, rp, m, t, cfg)
    mask = annulus_mask(xT)
    assert int(mask.sum()) > 500
    window_loss(xT, vT, mask).backward()

    def fwd(pmv, rpv, vv):
        x, v = run_window64(x0, vv, pmv, rpv, m, t, cfg)
        return window_loss(x, v, mask).item()

    h = 1e-4
    fd_m = (fwd(10.0 * (1 + h), 5.2, v0) - fwd(10.0 * (1 - h), 5.2, v0)) / (2 * h * 10.0)
    fd_r = (fwd(10.0, 5.2 * (1 + h), v0) - fwd(10.0, 5.2 * (1 - h), v0)) / (2 * h * 5.2)
    g = torch.Generator(device="cuda").manual_seed(10)
    dv = torch.randn(3, n, generator=g, device="cuda")
    # the velocities are ~1 in these units: h is also the relative step here
    fd_v = (fwd(10.0, 5.2, v0 + h * dv) - fwd(10.0, 5.2, v0 - h * dv)) / (2 * h)
    ad_v = (v0g.grad.double() * dv.double()).sum().item()
    print(f"  dm {pm.grad.item():.6e} vs {fd_m:.6e}; dr {rp.grad.item():.6e} vs {fd_r:.6e}; "
          f"dv {ad_v:.6e} vs {fd_v:.6e}", flush=True)
    assert abs(pm.grad.item() - fd_m) <= 3e-3 * abs(fd_m)
    assert abs(rp.grad.item() - fd_r) <= 3e-3 * abs(fd_r)
    assert abs(ad_v - fd_v) <= 3e-3 * abs(fd_v)


def test_window_checkpoint_interval_does_not_change_gradient(gpu):
    """Recomputation from checkpoints reproduces the forward bit for bit, so
    the gradient depends on the interval only through FP64 summation order."""
    nbx = gpu
    from nbx.window import window
    n, steps = 4096, 12
    out = []
    for ck in (1, 5, 12):
        d, x0, v0, m, t, cfg = window_case(nbx, n, 51, 0.5, steps, ck)
        print(f"checkpoint interval {ck}", flush=True)
        pm = torch.tensor(10.0, dtype=F64, requires_grad=True)
        rp = torch.tensor(5.2, dtype=F64, requires_grad=True)
        v0g = v0.clone().requires_grad_(True)
        xT, vT = window(x0, v0g, pm, rp, m, t, cfg)
        (xT.double().sum() + vT.double().pow(2).sum()).backward()
        out.append((xT.clone(), v0g.grad.double().clone(), pm.grad.item(), rp.grad.item()))
    for o in out[1:]:
        assert torch.equal(o[0], out[0][0])
        assert torch.allclose(o[1], out[0][1], rtol=1e-10, atol=0)
        assert abs(o[2] - out[0][2]) <= 1e-10 * abs(out[0][2])
        assert abs(o[3] - out[0][3]) <= 1e-10 * abs(out[0][3])


def test_reverse_mode_argument_errors(gpu):
    nbx = gpu
    n = 256
    d = disk(n, seed=5)
    x, v, m = dev_state(nbx, d)
    p = particles(nbx, x, v, torch.zeros_like(x), m)
    t = nbx.allocateOctree(n)
    dp = nbx.disk_params(0.0, None)
    ab = torch.zeros(3, n, dtype=F64, device="cuda")
    xb = torch.zeros_like(ab)
    L = nbx.lib()
    args = (C.byref(p.c()), 0.5, 1e-3, 1.0, C.byref(dp), 0.0, C.c_void_p(ab.data_ptr()),
            C.c_void_p(xb.data_ptr()), None)
    assert L.nbx_adj_force(t.h, *args) == -4  # reverse mode not enabled
    assert L.nbx_tree_set_adjoint(t.h, 1) == 0
    nbx.set_multipole(t, 2)
    assert L.nbx_adj_force(t.h, *args) == -1  # quadrupole walk has no adjoint
    nbx.set_multipole(t, 1)
    assert L.nbx_adj_force(t.h, *args) == 0
    sp = nbx.step_params(collisions=True)
    vb = torch.zeros_like(ab)
    rc = L.nbx_adj_step(t.h, C.byref(p.c()), C.byref(p.c()), C.byref(sp), C.byref(dp), 0,
                        C.c_void_p(xb.data_ptr()), C.c_void_p(vb.data_ptr()),
                        C.c_void_p(ab.data_ptr()), None)
    assert rc == -1 and b"collisions" in L.nbx_last_error()
    tb = nbx.allocateOctreeBatch([128, 128])
    assert L.nbx_tree_set_adjoint(tb.h, 1) == 0
    assert L.nbx_adj_force(tb.h, *args) == -1  # one disk per tree
