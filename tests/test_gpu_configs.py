"""GPU parity at the benchmark workloads' own sizes, through the C ABI.

One test per workload: the 4,096-body disk over 100 steps with FoF every 50
steps; the 131,072-body disk with collisions, drag and the 10 M_Jup
perturber stepped against the oracle's restated reference loop, with FoF at
the reference's planet-formation link length; the 1,048,576-body disk's walk
at theta 0.3 / 0.5 / 0.7 against the oracle's walk on 16,384 seeded targets;
the six-disk perturber sweep (16,384 ... 262,144 bodies) where every member,
stepped in one batched tree and in two concurrently stepped groups, must
equal its own plain run bit for bit.  Plus the step paths a large walk takes
(collision pass submitted before the walk), a caller's stream, and searches
below clumps deeper than any fixed stack.
"""
import ctypes as C

import numpy as np
import pytest
import torch

import pyoracle as O
from helpers import disk, rel_err

pytestmark = pytest.mark.gpu

STATE = ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az")
LL_PLANET = 0.0261  # link length of the reference's planet_formation.json run


def _np(t):
    return t.detach().cpu().numpy()


def _say(msg):
    print(msg, flush=True)


def _oracle_params(r_coll, pert_mjup, drag=1e-3, collisions=1):
    return O.step_params(dt=0.01, theta=0.5, softening=1e-3, star_mass=1.0, r_coll=r_coll,
                         restitution=0.0, gas_drag_rate=drag,
                         perturber_enabled=int(pert_mjup > 0), pert_mass_mjup=pert_mjup,
                         pert_radius_au=5.2, collisions=collisions, exact_collide=1)


def _gpu_params(nbx, use_graph, collisions=True):
    return nbx.step_params(dt=0.01, theta=0.5, softening=1e-3, star_mass=1.0, restitution=0.0,
                           gas_drag_rate=1e-3, collisions=collisions, use_graph=use_graph)


# ---------------------------------------------------------------------------
# 4,096 bodies, 100 steps, FoF every 50 (the small plain-method workload)
# ---------------------------------------------------------------------------
def test_small_disk_100_steps_with_fof(gpu):
    """100 KDK steps (drag 1e-3, collisions at r_coll = particlePhysicalRadius)
    against the oracle's restated reference loop: positions within 2e-5 AU
    and velocities within 5e-5 AU/T0 after 100 steps (measured 2.7e-6 and
    1.0e-5: the tree sum is added to the star term once here, inside the
    star accumulator there, so the states drift apart by a few ulp per
    step); FoF labels after steps 50 and 100 bit-exact against the
    oracle's partition of the same positions."""
    nbx = gpu
    n = 4096
    d = disk(n, mass=1e-5, seed=101)
    r_coll = nbx.particlePhysicalRadius(np.float32(1e-5 / n))
    sp = _gpu_params(nbx, use_graph=True)
    dp = nbx.disk_params(r_coll)
    p = nbx.to_gpu(d)
    t = nbx.allocateOctree(n)
    nbx.init_forces(t, p, sp, dp)
    s = O.ref_steps(d, d["mass"], _oracle_params(r_coll, 0.0), 0, 0, init_forces=True)
    for half in range(2):
        _say(f"n=4096 steps {50 * half}..{50 * half + 50}")
        nbx.step(t, p, sp, dp, 50 * half, 50)
        s = O.ref_steps(s, d["mass"], _oracle_params(r_coll, 0.0), 50 * half, 50)
        g = p.numpy()
        for ll in (r_coll, LL_PLANET):
            lab = _np(nbx.fof_labels(t, p, ll))
            ref, _ = O.fof(g["x"], g["y"], g["z"], ll)
            np.testing.assert_array_equal(lab, ref)
    err = {k: float(np.abs(g[k] - s[k]).max()) for k in ("x", "y", "z", "vx", "vy", "vz")}
    _say(f"max |GPU - oracle| after 100 steps: {err}")
    for k, e in err.items():
        assert e <= (2e-5 if k in ("x", "y", "z") else 5e-5), (k, e)


# ---------------------------------------------------------------------------
# 131,072 bodies: the benchmark step against the oracle loop
# ---------------------------------------------------------------------------
def test_bench_disk_steps_vs_oracle_and_fof(gpu):
    """Ten full steps of the benchmark disk (collisions, drag, 10 M_Jup
    perturber) against the oracle's restated reference loop with the exact
    pair set: positions and velocities within 4e-6; then FoF at the
    planet-formation link length (0.0261 AU, non-trivial groups) bit-exact
    against the oracle's grid partition of the same positions."""
    nbx = gpu
    n = 131072
    d = disk(n, mass=3e-5, seed=2024)
    r_coll = nbx.particlePhysicalRadius(np.float32(3e-5 / n))
    sp = _gpu_params(nbx, use_graph=True)
    dp = nbx.disk_params(r_coll, nbx.makePerturber(True, 10.0, 5.2, 0.0, 1.0))
    p = nbx.to_gpu(d)
    t = nbx.allocateOctree(n)
    _say("n=131072 10 steps")
    nbx.init_forces(t, p, sp, dp)
    nbx.step(t, p, sp, dp, 0, 10)
    g = p.numpy()
    s = O.ref_steps(d, d["mass"], _oracle_params(r_coll, 10.0), 0, 10, init_forces=True)
    err = {k: float(np.abs(g[k] - s[k]).max()) for k in ("x", "y", "z", "vx", "vy", "vz")}
    _say(f"max |GPU - oracle| after 10 steps: {err}")
    for k, e in err.items():
        assert e <= 4e-6, (k, e)
    _say("n=131072 FoF ll=0.0261")
    lab = _np(nbx.fof_labels(t, p, LL_PLANET))
    ref, groups = O.fof(g["x"], g["y"], g["z"], LL_PLANET)
    np.testing.assert_array_equal(lab, ref)
    assert (np.bincount(ref) >= 2).sum() > 1000  # thousands of multi-body groups


@pytest.mark.parametrize("shape", [16, 32])
def test_large_walk_step_paths_graph_equals_direct(gpu, shape):
    """At 131,072 bodies with 16 targets per wave the walk has more waves
    than the GPU holds at once, so the step submits the collision pass on
    its side stream before the walk; with 32 it submits the walk first.
    Either way the replayed step graph and direct launches give the same
    state bit for bit over five steps with collisions."""
    nbx = gpu
    n = 131072
    d = disk(n, mass=3e-5, seed=7)
    r_coll = np.float32(2e-3)  # many colliding pairs
    dp = nbx.disk_params(r_coll, nbx.makePerturber(True, 10.0, 5.2, 0.0, 1.0))
    out = []
    for use_graph in (True, False):
        _say(f"n=131072 shape={shape} graph={use_graph}")
        sp = _gpu_params(nbx, use_graph)
        p = nbx.to_gpu(d)
        t = nbx.allocateOctree(n)
        nbx.set_walk_shape(t, shape)
        nbx.init_forces(t, p, sp, dp)
        nbx.step(t, p, sp, dp, 0, 3)
        nbx.step(t, p, sp, dp, 3, 2)
        out.append(p.numpy())
    for k in STATE:
        np.testing.assert_array_equal(out[0][k], out[1][k], err_msg=k)


# ---------------------------------------------------------------------------
# 1,048,576 bodies: the walk at three opening angles
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("theta", [0.3, 0.5, 0.7])
def test_million_body_walk_vs_oracle(gpu, theta):
    """Self-gravity of the 2^20-body disk (r 0.3-10 AU) against the oracle's
    walk of the same compressed octree on 16,384 seeded targets: median rel
    diff <= 2e-6, 99.9th pct <= 1e-3 (COM rounding, hardware rsqrt)."""
    nbx = gpu
    n = 1 << 20
    d = disk(n, mass=3e-5, r_min=0.3, r_max=10.0, seed=99)
    p = nbx.to_gpu(d)
    t = nbx.allocateOctree(n)
    _say(f"n=2^20 theta={theta}")
    nbx.launchBBoxKernel(t, p)
    nbx.launchBuildKernel(t, p)
    nbx.launchSummarizeKernel(t)
    sa = [torch.zeros(n, dtype=torch.float32, device="cuda") for _ in range(3)]
    nbx.launchForceKernel(t, p, theta, 1e-3, 1.0, self_acc=sa)
    tg = np.sort(np.random.default_rng(1).choice(n, 16384, replace=False)).astype(np.int32)
    g = np.stack([_np(s)[tg] for s in sa], 1)
    mt = O.mt_force(d["x"], d["y"], d["z"], d["mass"], theta, 1e-3, targets=tg)
    e = rel_err(g, mt["self_a"])
    _say(f"theta={theta}: median {np.median(e):.3e}, p99.9 {np.percentile(e, 99.9):.3e}")
    assert np.median(e) <= 2e-6, np.median(e)
    assert np.percentile(e, 99.9) <= 1e-3, np.percentile(e, 99.9)


# ---------------------------------------------------------------------------
# The six-disk perturber sweep
# ---------------------------------------------------------------------------
SWEEP_N = [16384, 32768, 65536, 98304, 131072, 262144]
SWEEP_MJUP = [0.0, 0.3, 1.0, 3.0, 5.0, 10.0]


def test_sweep_members_equal_plain_runs(gpu):
    """The six sweep disks (uneven N, perturbers 0-10 M_Jup, collisions,
    drag) stepped 20 times three ways: each alone (its own tree), all six
    in one batched tree, and as two batched groups stepped side by side on
    their own streams (the sweep's grouping).  Every member equals its plain
    run bit for bit in all nine state arrays, and its FoF labels on the
    batched tree (each disk at its own link length, 0.005-0.0261 AU) equal
    the oracle's partition of its positions."""
    nbx = gpu
    disks = [disk(n, mass=3e-5, seed=100 + i) for i, n in enumerate(SWEEP_N)]
    sp = _gpu_params(nbx, use_graph=True)
    dps = []
    for n, mj in zip(SWEEP_N, SWEEP_MJUP):
        r_coll = nbx.particlePhysicalRadius(np.float32(3e-5 / n))
        dps.append(nbx.disk_params(r_coll, nbx.makePerturber(mj > 0, mj, 5.2, 0.0, 1.0)))
    plain = []
    for i, (dd, dp) in enumerate(zip(disks, dps)):
        _say(f"plain run of disk {i} (n={SWEEP_N[i]})")
        p = nbx.to_gpu(dd)
        t = nbx.allocateOctree(p.n)
        nbx.init_forces(t, p, sp, dp)
        nbx.step(t, p, sp, dp, 0, 20)
        plain.append(p.numpy())
        t.free()
        del p

    def check(g, label):
        off = 0
        for i, (n, ref) in enumerate(zip(SWEEP_N, plain)):
            for k in STATE:
                np.testing.assert_array_equal(g[k][off:off + n], ref[k],
                                              err_msg=f"{label}: disk {i} {k}")
            off += n

    _say("one batched tree of six disks")
    cat = {k: np.concatenate([dd[k] for dd in disks]) for k in disks[0]}
    pb = nbx.to_gpu(cat)
    tb = nbx.allocateOctreeBatch(SWEEP_N)
    nbx.init_forces(tb, pb, sp, dps)
    nbx.step(tb, pb, sp, dps, 0, 20)
    gb = pb.numpy()
    check(gb, "batched tree")
    _say("FoF of the batched tree, one link length per disk")
    lls = [LL_PLANET, 0.01, LL_PLANET, 0.005, 0.02, LL_PLANET]
    lab_t = torch.empty(pb.n, dtype=torch.int32, device="cuda")
    nbx.fof_into(tb, pb, lls, lab_t)
    lab = _np(lab_t)
    off = 0
    for i, (n, ll) in enumerate(zip(SWEEP_N, lls)):
        ref, _ = O.fof(gb["x"][off:off + n], gb["y"][off:off + n], gb["z"][off:off + n], ll)
        np.testing.assert_array_equal(lab[off:off + n] - off, ref, err_msg=f"FoF disk {i}")
        off += n
    tb.free()
    del pb

    _say("two groups stepped side by side")
    groups = nbx.split_groups(SWEEP_N, 2)
    trees, parts, gdps = [], [], []
    for g in groups:
        c = {k: np.concatenate([disks[i][k] for i in g]) for k in disks[0]}
        parts.append(nbx.to_gpu(c))
        trees.append(nbx.allocateOctreeBatch([SWEEP_N[i] for i in g]))
        gdps.append([dps[i] for i in g])
        nbx.init_forces(trees[-1], parts[-1], sp, gdps[-1])
    nbx.step_concurrent(trees, parts, sp, gdps, 0, 7)
    nbx.step_concurrent(trees, parts, sp, gdps, 7, 13)
    got = [q.numpy() for q in parts]
    cat_g = {k: np.concatenate([x[k] for x in got]) for k in STATE}
    assert sum(groups, []) == list(range(6))
    check(cat_g, "concurrent groups")


# ---------------------------------------------------------------------------
# Stream ownership and deep searches
# ---------------------------------------------------------------------------
def test_caller_stream_survives_tree(gpu):
    """A tree switched to a caller's stream works on it, returns to its own
    stream with NULL, and never destroys the caller's stream."""
    nbx = gpu
    L = nbx.lib()
    d = disk(5000, seed=3)
    p = nbx.to_gpu(d)
    s = torch.cuda.Stream()
    t = nbx.allocateOctree(p.n)
    assert L.nbx_tree_set_stream(t.h, C.c_void_p(s.cuda_stream)) == 0
    nbx.launchBBoxKernel(t, p)
    nbx.launchBuildKernel(t, p)
    nbx.launchSummarizeKernel(t)
    nbx.launchForceKernel(t, p, 0.5, 1e-3, 1.0)
    a1 = _np(p.ax)
    assert L.nbx_tree_set_stream(t.h, None) == 0
    nbx.launchForceKernel(t, p, 0.5, 1e-3, 1.0)
    np.testing.assert_array_equal(_np(p.ax), a1)
    assert L.nbx_tree_set_stream(t.h, C.c_void_p(s.cuda_stream)) == 0
    t.free()
    with torch.cuda.stream(s):  # the caller's stream is still usable
        x = torch.arange(1000, device="cuda", dtype=torch.float32) * 2
    s.synchronize()
    assert float(x.sum()) == 999000.0


def _deep_clump(n_bg, n_clump, seed):
    """A disk plus a clump whose bodies sit at log-uniform distances
    3e-6 ... 5e-3 AU from one point: the octree below the clump spans ~11
    levels with bodies in several octants of every level, so a depth-first
    search with every cell in range keeps dozens of pending blocks."""
    d = disk(n_bg, seed=seed)
    rng = np.random.default_rng(seed)
    r = 10.0 ** rng.uniform(np.log10(3e-6), np.log10(5e-3), n_clump)
    u = rng.normal(size=(n_clump, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    c = np.array([1.3, 0.002, 0.7]) + r[:, None] * u
    f = np.float32
    for k, col in zip(("x", "y", "z"), c.T):
        d[k] = np.concatenate([d[k], col.astype(f)])
    for k in ("vx", "vy", "vz"):
        d[k] = np.concatenate([d[k], rng.normal(0, 0.05, n_clump).astype(f)])
    for k in ("ax", "ay", "az"):
        d[k] = np.concatenate([d[k], np.zeros(n_clump, f)])
    d["mass"] = np.concatenate([d["mass"], np.full(n_clump, d["mass"][0], f)])
    d["id"] = np.arange(len(d["x"]), dtype=np.int32)
    return d


@pytest.mark.parametrize("ll", [1e-2, 1e-4, 1e-6])
def test_deep_clump_fof_and_collisions(gpu, ll):
    """FoF and collision searches below a clump of 3,000 bodies spread over
    three decades of scale (every cell of the clump within the search radius
    at ll = 1e-2, where the pending blocks overflow the per-lane stack and
    the search goes on without it; partial at 1e-4 and 1e-6): the pair count
    equal to the oracle's, labels bit-exact against the
    oracle, collision deltas against the oracle's exact pair set within
    1e-5 max |v| (a clump body sums up to 3,000 impulses, in Morton
    pre-order here and in the reference's stack order there; measured
    1.9e-6)."""
    nbx = gpu
    d = _deep_clump(20000, 3000, seed=5)
    n = len(d["x"])
    p = nbx.to_gpu(d)
    t = nbx.allocateOctree(n)
    _say(f"deep clump ll={ll}")
    nbx.search_stats(t, clear=True)
    lab = _np(nbx.fof_labels(t, p, ll))
    unstacked = nbx.search_stats(t)["unstacked"]
    if ll == 1e-2:  # pending blocks pile up past the per-lane stack
        assert unstacked > 0
    ref, _ = O.fof(d["x"], d["y"], d["z"], ll, method="grid")
    np.testing.assert_array_equal(lab, ref)
    if ll == 1e-2:
        assert np.bincount(ref).max() >= 3000
    nbx.launchBBoxKernel(t, p)
    nbx.launchBuildKernel(t, p)
    nbx.launchSummarizeKernel(t)
    nbx.search_stats(t, clear=True)
    nbx.launchCollisionKernel(t, p, ll, 0.2)
    g = p.numpy()
    rt = O.RefTree(d["x"], d["y"], d["z"], d["mass"])
    assert rt.lost == 0  # the oracle's canonical octree holds every body
    vx, vy, vz, pairs = rt.collide(d["vx"], d["vy"], d["vz"], ll, 0.2, exact=True)
    # every ordered pair counted once by each of its bodies
    assert nbx.search_stats(t)["pairs"] == pairs
    gv = np.stack([g["vx"], g["vy"], g["vz"]], 1)
    ov = np.stack([vx, vy, vz], 1)
    assert np.abs(gv - ov).max() <= 1e-5 * np.abs(ov).max()
    assert pairs > 0


def test_wave_orders_give_identical_results(gpu):
    """The launch order of the walk's wavefronts (Morton order, longest first,
    contiguous ranges per XCD) changes only the schedule: forces and ten
    stepped states are bit-identical across the three orders."""
    nbx = gpu
    n = 65536
    d = disk(n, mass=3e-5, seed=88)
    sp = _gpu_params(nbx, use_graph=True)
    dp = nbx.disk_params(np.float32(2e-3), nbx.makePerturber(True, 10.0, 5.2, 0.0, 1.0))
    out = []
    for order in (nbx.WAVES_LAUNCH, nbx.WAVES_LONGEST_FIRST, nbx.WAVES_XCD):
        _say(f"wave order {order}")
        p = nbx.to_gpu(d)
        t = nbx.allocateOctree(n)
        nbx.set_wave_order(t, order)
        nbx.init_forces(t, p, sp, dp)
        nbx.step(t, p, sp, dp, 0, 10)
        out.append(p.numpy())
    for o in out[1:]:
        for k in STATE:
            np.testing.assert_array_equal(o[k], out[0][k], err_msg=k)
