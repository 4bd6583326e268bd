"""The FP64 step (NBX_PRECISION_FP64), the reference of the precision study.

It has no counterpart in the reference, so it is pinned from four sides:
- at theta = 0 the walk accepts leaves only, so its accelerations must equal
  the oracle's FP64 direct sum (itself pinned by brute force and the
  two-body closed form) to FP64 rounding;
- stepped with drag and the perturber, it must stay within float tolerance of
  the oracle's restated float step loop (main.cu:85-140): same terms, same
  signs, same orbit of the perturber;
- KDK is time-reversible: k steps with dt then k steps with -dt return to the
  start exactly in exact arithmetic, so the FP64 step returns to ~1e-13 while
  any float rounding on the state path leaves ~1e-7 (the float step is run
  beside it to show the test tells them apart);
- a batched tree's members equal their plain FP64 runs bit for bit, and the
  exported state and the FP64 diagnostics read the same numbers.
"""
import numpy as np
import pytest
import torch

import pyoracle as O
from helpers import disk

pytestmark = pytest.mark.gpu

STATE = ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az")


def _say(msg):
    print(msg, flush=True)


def _s64(nbx, t, p):
    return nbx.state64(t, p).cpu().numpy()


def _params(nbx, theta=0.5, star_mass=1.0, drag=0.0, dt=0.01):
    return nbx.step_params(dt=dt, theta=theta, softening=1e-3, star_mass=star_mass,
                           restitution=0.0, gas_drag_rate=drag, collisions=False)


def test_fp64_walk_at_theta0_is_the_fp64_direct_sum(gpu):
    """theta = 0, no star, no perturber: the FP64 walk's accelerations equal
    the oracle's FP64 direct sum of the same positions to FP64 rounding
    (relative per body: median <= 1e-13, max <= 1e-10, where a body's
    neighbours nearly cancel); the accelerations handed back in the float
    arrays are those values rounded."""
    nbx = gpu
    n = 2048
    d = disk(n, mass=1e-5, seed=41)
    _say(f"fp64 walk theta=0 n={n}")
    p = nbx.to_gpu(d)
    t = nbx.allocateOctree(n)
    nbx.set_precision(t, "fp64")
    sp = _params(nbx, theta=0.0, star_mass=0.0)
    nbx.init_forces(t, p, sp, nbx.disk_params(0.0, None))
    a64 = _s64(nbx, t, p)[6:9].T
    ref = O.direct_f64(d["x"], d["y"], d["z"], d["mass"], float(np.float32(1e-3)))
    err = np.linalg.norm(a64 - ref, axis=1) / np.linalg.norm(ref, axis=1)
    assert np.median(err) <= 1e-13 and err.max() <= 1e-10, (np.median(err), err.max())
    # the float accelerations handed to the caller are the rounded FP64 ones
    g = p.numpy()
    for k, c in zip(("ax", "ay", "az"), range(3)):
        assert np.array_equal(g[k], a64[:, c].astype(np.float32)), k


def _oracle_loop(d, drag, pert, steps=20):
    params = O.step_params(dt=0.01, theta=0.5, softening=1e-3, star_mass=1.0, r_coll=0.0,
                           restitution=0.0, gas_drag_rate=drag, perturber_enabled=int(pert),
                           pert_mass_mjup=10.0, pert_radius_au=5.2, collisions=0,
                           exact_collide=1)
    return O.ref_steps(d, d["mass"], params, 0, steps, init_forces=True)


def _max_diff(a, b):
    return max(np.abs(a[k] - b[k]).max() for k in ("x", "y", "z", "vx", "vy", "vz"))


def test_fp64_step_matches_reference_loop(gpu):
    """20 steps at N = 4,096 with the 10 M_Jup perturber, against the oracle's
    restated float loop (collisions off): drag 1e-3 (the workload's) within
    4e-6, and a strong drag (rate 1) within 1e-5, where leaving out the drag
    or the perturber would move the state by more than 1e-4: the FP64 step
    has the reference's terms and signs and differs from the float loop only
    by float rounding."""
    nbx = gpu
    d = disk(4096, mass=1e-5, seed=9)
    pert = nbx.makePerturber(True, 10.0, 5.2, 0.0, 1.0)
    for drag, tol in ((1e-3, 4e-6), (1.0, 1e-5)):
        _say(f"fp64 step vs oracle loop drag={drag}")
        sp = _params(nbx, drag=drag)
        dp = nbx.disk_params(0.0, pert)
        p = nbx.to_gpu(d)
        t = nbx.allocateOctree(p.n)
        nbx.set_precision(t, "fp64")
        nbx.init_forces(t, p, sp, dp)
        nbx.step(t, p, sp, dp, 0, 20)
        g = _s64(nbx, t, p)
        s = _oracle_loop(d, drag, True)
        for c, k in enumerate(("x", "y", "z", "vx", "vy", "vz")):
            err = np.abs(g[c] - s[k]).max()
            assert err <= tol, (drag, k, err)
    # the strong-drag case would see a missing term
    assert _max_diff(s, _oracle_loop(d, 0.0, True)) > 1e-4
    assert _max_diff(s, _oracle_loop(d, 1.0, False)) > 1e-4


def test_fp64_step_is_time_reversible(gpu):
    """50 steps with dt = 0.01 then 50 with dt = -0.01 (star, self-gravity at
    theta = 0, no drag, no perturber) return the FP64 step to its start
    within 1e-11 AU and 1e-11 AU/T0; the float step, run the same way,
    misses by more than 1e-8."""
    nbx = gpu
    n = 1024
    d = disk(n, mass=1e-5, seed=5)
    miss = {}
    for mode in ("fp64", "float"):
        _say(f"time reversal mode={mode}")
        p = nbx.to_gpu(d)
        t = nbx.allocateOctree(n)
        nbx.set_precision(t, mode)
        fwd, bwd = _params(nbx, theta=0.0), _params(nbx, theta=0.0, dt=-0.01)
        dp = nbx.disk_params(0.0, None)
        nbx.init_forces(t, p, fwd, dp)
        nbx.step(t, p, fwd, dp, 0, 50)
        nbx.step(t, p, bwd, dp, 0, 50)
        g = _s64(nbx, t, p)
        miss[mode] = max(np.abs(g[c] - d[k].astype(np.float64)).max()
                         for c, k in enumerate(("x", "y", "z", "vx", "vy", "vz")))
    assert miss["fp64"] <= 1e-11, miss
    assert miss["float"] > 1e-8, miss


def test_fp64_batched_members_equal_plain_runs(gpu):
    """Two disks of uneven N in one batched FP64 tree, 5 steps with drag and
    the perturber on one of them: each member bit-identical to its plain FP64
    run."""
    nbx = gpu
    counts = [3000, 1024]
    disks = [disk(n, mass=1e-5, seed=60 + i) for i, n in enumerate(counts)]
    sp = _params(nbx, drag=1e-3)
    dps = [nbx.disk_params(0.0, None),
           nbx.disk_params(0.0, nbx.makePerturber(True, 10.0, 5.2))]
    plain = []
    for dd, dp in zip(disks, dps):
        _say(f"fp64 plain run n={len(dd['x'])}")
        p = nbx.to_gpu(dd)
        t = nbx.allocateOctree(p.n)
        nbx.set_precision(t, "fp64")
        nbx.init_forces(t, p, sp, dp)
        nbx.step(t, p, sp, dp, 0, 5)
        plain.append(_s64(nbx, t, p))
    _say("fp64 batched run")
    cat = {k: np.concatenate([dd[k] for dd in disks]) for k in disks[0]}
    p = nbx.to_gpu(cat)
    t = nbx.allocateOctreeBatch(counts)
    nbx.set_precision(t, "fp64")
    nbx.init_forces(t, p, sp, dps)
    nbx.step(t, p, sp, dps, 0, 5)
    g = _s64(nbx, t, p)
    off = 0
    for n, ref in zip(counts, plain):
        assert np.array_equal(g[:, off:off + n], ref), n
        off += n


def test_state_export_and_diagnostics_per_mode(gpu):
    """state64 is the FP64 state (FP64 mode), hi + lo (compensated: lo is
    nonzero after steps) or the float state; the FP64 diagnostics of each
    mode equal the kinetic energy, star potential and angular momentum
    summed from that exported state in numpy (relative 1e-12)."""
    nbx = gpu
    n = 4096
    d = disk(n, mass=1e-5, seed=3)
    sp = _params(nbx, drag=1e-3)
    dp = nbx.disk_params(0.0, nbx.makePerturber(True, 10.0, 5.2))
    m = d["mass"].astype(np.float64)
    for mode in ("float", "compensated", "fp64"):
        _say(f"state export mode={mode}")
        p = nbx.to_gpu(d)
        t = nbx.allocateOctree(n)
        nbx.set_precision(t, mode)
        nbx.init_forces(t, p, sp, dp)
        nbx.step(t, p, sp, dp, 0, 10)
        g = _s64(nbx, t, p)
        f = p.numpy()
        hi = np.stack([f[k] for k in STATE]).astype(np.float64)
        if mode == "float":
            assert np.array_equal(g, hi)
        else:
            assert np.array_equal(g[:6].astype(np.float32), hi[:6].astype(np.float32))
            assert np.abs(g[:6] - hi[:6]).max() > 0.0  # the state carries more than floats
        x, y, z, u, v, w = g[:6]
        kin = 0.5 * np.sum(m * (u * u + v * v + w * w))
        eps = float(np.float32(1e-3))
        pot = -np.sum(m / np.sqrt(x * x + y * y + z * z + eps * eps))
        ly = np.sum(m * (z * u - x * w))
        dg = nbx.diagnostics(t, p, 1.0, 1e-3)
        assert abs(dg["kinetic"] - kin) <= 1e-12 * abs(kin), (mode, dg["kinetic"], kin)
        assert abs(dg["star_potential"] - pot) <= 1e-12 * abs(pot), mode
        assert abs(dg["L"][1] - ly) <= 1e-12 * abs(ly), mode


def test_fp64_mode_rejects_what_it_leaves_out(gpu):
    """Collisions or quadrupoles in FP64 mode, and a step before the FP64
    state is loaded, are errors."""
    nbx = gpu
    d = disk(1024, mass=1e-5, seed=2)
    p = nbx.to_gpu(d)
    t = nbx.allocateOctree(p.n)
    nbx.set_precision(t, "fp64")
    dp = nbx.disk_params(0.0, None)
    with pytest.raises(RuntimeError, match="FP64 state not loaded"):
        nbx.step(t, p, _params(nbx), dp, 0, 1)
    coll = nbx.step_params(dt=0.01, theta=0.5, softening=1e-3, restitution=0.0,
                           gas_drag_rate=0.0, collisions=True)
    with pytest.raises(RuntimeError, match="no collision pass"):
        nbx.init_forces(t, p, coll, dp)
    nbx.set_multipole(t, 2)
    with pytest.raises(RuntimeError, match="monopoles only"):
        nbx.init_forces(t, p, _params(nbx), dp)
    nbx.set_multipole(t, 1)
    nbx.init_forces(t, p, _params(nbx), dp)
    nbx.step(t, p, _params(nbx), dp, 0, 1)
    with pytest.raises(RuntimeError):
        nbx.set_precision(t, 3)
