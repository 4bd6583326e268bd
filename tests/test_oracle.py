"""CPU tests: the oracle pinned before it is trusted.

Each oracle function is checked against something independent of its own
formula: the reference's known-answer tests (tests/*.cu, *.cpp of the
reference, cited per test), the reference's own FoF / disk code compiled
into oracle/_ref/libref.so, or what the mathematics fixes (brute force,
closed forms, invariants).
"""
import math
import os

import numpy as np
import pytest

import pyoracle as O
from helpers import clustered, disk, rel_err

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

f32 = np.float32


# ---------------------------------------------------------------------------
# Octree, order and walk (BHTree.cu restatement)
# ---------------------------------------------------------------------------
def _line10():
    d = {k: np.zeros(10, f32) for k in ("x", "y", "z", "mass")}
    d["x"] = (np.arange(10) * 0.5).astype(f32)
    d["mass"][:] = 3.003e-7
    return d


def test_ref_sort_is_permutation():
    """test_bhtree.cu:29-99: 10 bodies on a line and the 1000-body ring."""
    d = _line10()
    t = O.RefTree(d["x"], d["y"], d["z"], d["mass"])
    assert sorted(t.sort_idx().tolist()) == list(range(10))
    n = 1000
    i = np.arange(n)
    r = 0.5 + 4.5 * i / (n - 1)
    phi = 2 * np.pi * i / n
    x, z = (r * np.cos(phi)).astype(f32), (r * np.sin(phi)).astype(f32)
    y = (0.01 * (i % 10 - 5)).astype(f32)
    t = O.RefTree(x, y, z, np.full(n, 3.003e-7, f32))
    assert t.lost == 0
    assert sorted(t.sort_idx().tolist()) == list(range(n))


def test_ref_single_particle_force_kat():
    """test_bhtree.cu:104-134: a_x = -1 within 5 % for a body at (1,0,0)."""
    one = np.ones(1, f32)
    t = O.RefTree(one, 0 * one, 0 * one, 1e-6 * one)
    a = t.force(0.5, 0.001, 1.0)["a"][0]
    assert abs(a[0] + 1.0) <= 0.05 and abs(a[1]) <= 0.05 and abs(a[2]) <= 0.05


def test_morton_order_equals_reference_dfs():
    """The DFS of the canonical octree (BHTree.cu:322-352) visits leaves in
    the order of the 21-digit keys of the same float descent; equal keys
    (none here) would be ordered by index instead."""
    for n, seed in [(257, 1), (4096, 2), (20000, 3)]:
        d = disk(n, seed=seed)
        keys = O.mt_keys(d["x"], d["y"], d["z"])
        assert len(np.unique(keys)) == n
        ref = O.RefTree(d["x"], d["y"], d["z"], d["mass"]).sort_idx()
        np.testing.assert_array_equal(O.mt_sort(keys), ref)


def test_keys_digits_match_octant_rule():
    """First digit = (x>cx) | (y>cy)<<1 | (z>cz)<<2 about the root centre
    (BHTree.cu:164-170), and keys never use bit 63."""
    d = disk(5000, seed=4)
    root = O.mt_root(d["x"], d["y"], d["z"])
    keys = O.mt_keys(d["x"], d["y"], d["z"], root)
    first = (keys >> np.uint64(60)).astype(int)
    want = ((d["x"] > root[0]).astype(int) | ((d["y"] > root[1]).astype(int) << 1)
            | ((d["z"] > root[2]).astype(int) << 2))
    np.testing.assert_array_equal(first, want)
    assert not np.any(keys >> np.uint64(63))


def test_root_cube_contains_all_bodies():
    d = disk(3000, seed=5)
    c = O.mt_root(d["x"], d["y"], d["z"])
    for k, ax in enumerate(("x", "y", "z")):
        assert np.all(np.abs(d[ax] - c[k]) < c[3])


def test_walk_theta_zero_is_direct_sum():
    """With a vanishing opening angle every leaf is accepted individually, so
    both walks reduce to the pairwise sum (float vs FP64 direct sum)."""
    d = disk(600, seed=6)
    ad = O.direct_f64(d["x"], d["y"], d["z"], d["mass"], 1e-3)
    ref = O.RefTree(d["x"], d["y"], d["z"], d["mass"]).force(1e-6, 1e-3, 1.0, counts=True)
    mt = O.mt_force(d["x"], d["y"], d["z"], d["mass"], 1e-6, 1e-3, counts=True)
    assert np.all(ref["n_acc"] == 599) and np.all(mt["n_acc"] == 599)
    assert rel_err(ref["self_a"], ad).max() < 1e-4
    assert rel_err(mt["self_a"], ad).max() < 1e-4


def test_compressed_walk_matches_canonical_walk():
    """The compressed octree accepts the same nodes as the canonical one
    (a folded single-child chain has the same COM and a larger cell), so the
    accepted counts agree and the forces agree to COM rounding."""
    for n, seed in [(4096, 7), (30000, 8)]:
        d = disk(n, seed=seed)
        ref = O.RefTree(d["x"], d["y"], d["z"], d["mass"]).force(0.5, 1e-3, 1.0, counts=True)
        mt = O.mt_force(d["x"], d["y"], d["z"], d["mass"], 0.5, 1e-3, counts=True)
        agree = np.mean(ref["n_acc"] == mt["n_acc"])
        assert agree > 0.999, agree
        assert np.all(mt["n_open"] <= ref["n_open"])
        e = rel_err(mt["self_a"], ref["self_a"])
        assert np.median(e) < 1e-5 and np.percentile(e, 99.9) < 1e-2


def test_quadrupole_improves_accuracy():
    d = disk(4096, seed=9)
    ad = O.direct_f64(d["x"], d["y"], d["z"], d["mass"], 1e-3)
    mono = O.mt_force(d["x"], d["y"], d["z"], d["mass"], 0.5, 1e-3)["self_a"]
    quad = O.mt_force(d["x"], d["y"], d["z"], d["mass"], 0.5, 1e-3, quad=True)["self_a"]
    assert np.median(rel_err(quad, ad)) < 0.5 * np.median(rel_err(mono, ad))


def test_direct_sum_matches_brute_force_and_two_body():
    rng = np.random.default_rng(1)
    n = 50
    x, y, z = (rng.normal(0, 1, n).astype(f32) for _ in range(3))
    m = rng.uniform(1e-7, 1e-6, n).astype(f32)
    eps = 1e-2
    a = O.direct_f64(x, y, z, m, eps)
    P = np.stack([x, y, z], 1).astype(np.float64)
    D = P[None, :, :] - P[:, None, :]
    r2 = (D ** 2).sum(-1) + eps * eps
    W = m[None, :] / (r2 * np.sqrt(r2))
    np.fill_diagonal(W, 0.0)
    np.testing.assert_allclose(a, (D * W[..., None]).sum(1), rtol=1e-12)
    # two bodies: |a_0| = m_1 d / (d^2+eps^2)^{3/2}, equal and opposite momentum
    x2 = np.array([0, 0.3], f32)
    a2 = O.direct_f64(x2, 0 * x2, 0 * x2, np.array([2e-6, 1e-6], f32), 0.0)
    d = float(x2[1])
    m0, m1 = float(f32(2e-6)), float(f32(1e-6))
    assert abs(a2[0, 0] - m1 / d ** 2) < 1e-15
    assert abs(m0 * a2[0, 0] + m1 * a2[1, 0]) < 1e-20


def test_overflow_and_loss_reported_not_silent():
    """Coincident bodies cannot be separated by the canonical build: the
    restatement reports them as lost (finding F6) rather than dropping them
    silently."""
    x = np.zeros(4, f32)
    x[2:] = 1.0
    t = O.RefTree(x, 0 * x, 0 * x, np.ones(4, f32) * 1e-6)
    assert t.lost > 0


# ---------------------------------------------------------------------------
# Collisions (BHTree.cu:620-708)
# ---------------------------------------------------------------------------
def test_collision_head_on_kat():
    """test_collision.cu:7-48: velocities reverse, momentum conserved."""
    x = np.array([-0.001, 0.001], f32)
    zero = np.zeros(2, f32)
    t = O.RefTree(x, zero, zero, np.full(2, 3.003e-7, f32))
    vx, vy, vz, pairs = t.collide(np.array([0.1, -0.1], f32), zero, zero, 0.005, 0.3)
    assert vx[0] < 0 and vx[1] > 0 and abs(vx[0] + vx[1]) <= 1e-5
    assert pairs == 2
    # closed form: equal masses, restitution e: v' = -e v along the normal
    assert abs(vx[0] + 0.3 * 0.1) < 1e-6


def test_collision_exact_pairs_superset_of_reference_pruning():
    """Finding F7: the reference prunes with sqrt(3) h around the COM and can
    miss pairs; the exact pass finds every pair of the brute-force set."""
    d = clustered(200, 4, 2e-3, seed=3)
    t = O.RefTree(d["x"], d["y"], d["z"], d["mass"])
    rc = 3e-3
    *_, p_exact = t.collide(d["vx"], d["vy"], d["vz"], rc, 0.0, exact=True)
    *_, p_ref = t.collide(d["vx"], d["vy"], d["vz"], rc, 0.0, exact=False)
    P = np.stack([d["x"], d["y"], d["z"]], 1)
    diff = P[:, None, :] - P[None, :, :]
    d2 = (diff.astype(f32) ** 2).sum(-1)
    brute = int((d2 < f32(rc) * f32(rc)).sum() - len(P))
    assert p_exact == brute
    assert p_ref <= p_exact


# ---------------------------------------------------------------------------
# Integrator, drag, perturber (Integrator.cu, Perturber.*)
# ---------------------------------------------------------------------------
def test_leapfrog_energy_kat():
    """test_integrator.cu:10-59: one orbit at r=1 conserves energy to 1 %."""
    x, v = [np.ones(1, f32), np.zeros(1, f32), np.zeros(1, f32)], [np.zeros(1, f32)] * 2
    v = [np.zeros(1, f32), np.zeros(1, f32), np.ones(1, f32)]
    dt, eps = f32(0.01), f32(0.001)

    def acc(x):
        r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2] + eps * eps
        r3 = r2 * np.sqrt(r2)
        return [-x[0] / r3, -x[1] / r3, -x[2] / r3]

    a = acc(x)
    for _ in range(int(f32(6.28318530) / dt)):
        v = O.half_kick(v, a, dt * f32(0.5))
        x = O.drift(x, v, dt)
        a = acc(x)
        v = O.half_kick(v, a, dt * f32(0.5))
    r = math.sqrt(sum(float(c[0]) ** 2 for c in x))
    e = 0.5 * sum(float(c[0]) ** 2 for c in v) - 1.0 / r
    assert abs(e + 0.5) <= 0.01 * 0.5


def test_kick_drift_are_float_expressions():
    rng = np.random.default_rng(2)
    v = [rng.normal(0, 1, 100).astype(f32) for _ in range(3)]
    a = [rng.normal(0, 1, 100).astype(f32) for _ in range(3)]
    out = O.half_kick(v, a, f32(0.005))
    for k in range(3):
        np.testing.assert_array_equal(out[k], v[k] + a[k] * f32(0.005))


def test_gas_drag_closed_forms():
    """Integrator.cu:40-76: circular Keplerian orbits (either sense) are fixed
    points; vertical and radial velocity decay by exp(-rate dt); the excess
    tangential speed relaxes toward v_kep; bodies on the axis are untouched;
    a non-positive rate is a no-op (Integrator.cu:102)."""
    r = f32(2.0)
    vk = float(np.sqrt(f32(1.0) / r))
    # bodies at (r,0,0): prograde tangential direction is +z
    xs = [(r, 0, 0)] * 4 + [(1e-7, 0, 1e-7)]
    vs = [(0, 0, vk), (0, 0, -vk), (0, 0.3, vk), (0.2, 0, 1.1 * vk), (0, 0.3, 0)]
    x = [np.array([p[k] for p in xs], f32) for k in range(3)]
    v = [np.array([q[k] for q in vs], f32) for k in range(3)]
    rate, dt = f32(0.5), f32(0.01)
    out = O.gas_drag(x, v, rate, dt, 1.0)
    decay = math.exp(-0.5 * 0.01)
    got = np.stack(out, 1).astype(np.float64)
    np.testing.assert_allclose(got[0], [0, 0, vk], atol=1e-6)
    np.testing.assert_allclose(got[1], [0, 0, -vk], atol=1e-6)
    np.testing.assert_allclose(got[2], [0, 0.3 * decay, vk], atol=1e-6)
    np.testing.assert_allclose(got[3], [0.2 * decay, 0, vk + 0.1 * vk * decay], atol=1e-6)
    np.testing.assert_array_equal(got[4], np.array(vs[4], f32))
    off = O.gas_drag(x, v, 0.0, dt, 1.0)
    for k in range(3):
        np.testing.assert_array_equal(off[k], v[k])


def test_perturber_orbit_kats():
    """test_perturber.cu:17-76."""
    p = O.make_perturber(1, 1.0, 10.0, 0.0, 1.0)
    assert p.enabled
    assert abs(p.x - 10.0) <= 1e-5 and abs(p.z) <= 1e-5 and abs(p.y) <= 1e-5
    assert abs(p.mass_msun - 9.5479e-4) <= 1e-4 * 9.5479e-4
    assert abs(p.omega - math.sqrt(1 / 1000)) <= 1e-5 * math.sqrt(1 / 1000)
    T = f32(2.0) * f32(3.14159265358979) / f32(p.omega)
    for frac, want in [(1.0, (10, 0)), (0.5, (-10, 0)), (0.25, (0, 10))]:
        q = O.make_perturber(1, 1.0, 10.0, 0.0, 1.0)
        O.update_perturber(q, float(T * f32(frac)))
        assert abs(q.x - want[0]) <= 1e-2 and abs(q.z - want[1]) <= 1e-2
    off = O.make_perturber(0, 1.0, 10.0, 0.0, 1.0)
    off.x = off.y = off.z = 42.0
    O.update_perturber(off, 1000.0)
    assert off.x == 42.0 and off.y == 42.0 and off.z == 42.0


def test_perturber_force_kats():
    """test_perturber.cu:83-149: +x pull on an origin body, '+=' semantics."""
    z = np.zeros(1, f32)
    a = O.perturber_force([z, z, z], [z, z, z], (10.0, 0.0, 0.0), 1e-3, 0.01)
    r2 = f32(100) + f32(0.01) * f32(0.01)
    want = f32(1e-3) / (r2 * np.sqrt(r2)) * f32(10)
    assert abs(a[0][0] - want) <= 1e-4 * want and abs(a[1][0]) <= 1e-7
    a = O.perturber_force([np.full(1, 1e-3, f32), z, z], [z, z, z], (10.0, 0, 0), 1e-3, 0.01)
    assert abs(a[0][0] - (1e-3 + want)) <= 1e-4 * (1e-3 + want)


def test_particle_radius_kat():
    """test_disk_init.cpp:42-56: r = 2.416e-5 AU at m = 3.003e-7 (1 %)."""
    r = O.particle_physical_radius(3.003e-7)
    assert abs(r - 2.416e-5) <= 0.01 * 2.416e-5
    ref = O.ref_lib()
    if ref is not None:
        assert r == ref.ref_particle_physical_radius(3.003e-7)


# ---------------------------------------------------------------------------
# FoF (Aggregate.cpp)
# ---------------------------------------------------------------------------
def test_fof_kats():
    """test_aggregate.cpp:20-119 on the oracle partition."""
    def lab(pos, ll):
        p = np.array(pos, f32)
        return O.fof(p[:, 0], p[:, 1], p[:, 2], ll)

    labels, g = lab([(0, 0, 0), (1e-4, 0, 0), (0, 1e-4, 0), (1e-4, 1e-4, 0), (0, 0, 1e-4),
                     (3, 0, 0), (-4, 0, 0)], 0.002)
    assert g == 1 and np.bincount(labels).max() == 5
    assert lab([(0, 0, 0), (10, 0, 0), (20, 0, 0)], 0.01)[1] == 0
    labels, g = lab([(0, 0, 0), (1e-3, 0, 0), (0, 1e-3, 0), (10, 0, 0), (10.001, 0, 0),
                     (10, 1e-3, 0)], 0.01)
    assert g == 2 and sorted(np.bincount(labels)[np.bincount(labels) > 0]) == [3, 3]


def test_fof_grid_equals_n2_and_brute_components():
    d = clustered(150, 5, 1e-3, seed=4)
    a, ga = O.fof(d["x"], d["y"], d["z"], 2e-3, method="n2")
    b, gb = O.fof(d["x"], d["y"], d["z"], 2e-3, method="grid")
    np.testing.assert_array_equal(a, b)
    assert ga == gb
    # connected components by breadth-first search on the float predicate
    P = np.stack([d["x"], d["y"], d["z"]], 1)
    n = len(P)
    adj = ((P[:, None, :] - P[None, :, :]) ** 2).sum(-1) < f32(2e-3) * f32(2e-3)
    comp = -np.ones(n, int)
    for s in range(n):
        if comp[s] >= 0:
            continue
        stack, comp[s] = [s], s
        while stack:
            u = stack.pop()
            for v in np.flatnonzero(adj[u]):
                if comp[v] < 0:
                    comp[v] = s
                    stack.append(v)
    np.testing.assert_array_equal(a, comp)


def test_fof_matches_reference_library():
    """oracle/_ref/libref.so is the reference's own Aggregate.cpp: same
    partition, and the host aggregate sums of nbx.aggregates_from_labels
    equal the reference's float sums bit for bit."""
    ref = O.ref_lib()
    if ref is None:
        pytest.skip("oracle/_ref/libref.so not built (needs the reference sources)")
    import nbx
    d = clustered(120, 4, 1e-3, seed=5)
    n = len(d["x"])
    labels = np.zeros(n, np.int32)
    cap = n
    agg_n = np.zeros(cap, np.int32)
    agg_f = np.zeros(cap * 8, np.float32)
    na = ref.ref_detect_aggregates(n, d["x"], d["y"], d["z"], d["vx"], d["vy"], d["vz"],
                                   2e-3, 3.003e-7, labels, cap, agg_n, agg_f)
    mine, groups = O.fof(d["x"], d["y"], d["z"], 2e-3)
    np.testing.assert_array_equal(mine, labels)
    assert groups == na
    aggs = nbx.aggregates_from_labels(mine, d, 3.003e-7)
    assert len(aggs) == na
    F = agg_f.reshape(cap, 8)
    for k, a in enumerate(aggs):
        assert a.n_particles == agg_n[k]
        got = [a.mass_msun, a.cx, a.cy, a.cz, a.vx, a.vy, a.vz, a.dist_center]
        assert [np.float32(g) for g in got] == [np.float32(v) for v in F[k]]


# ---------------------------------------------------------------------------
# The reference step loop restated (main.cu:85-140)
# ---------------------------------------------------------------------------
def test_reference_loop_two_body_kepler():
    """A single light body around the star (no tree partner) follows the
    leapfrog Kepler orbit: energy within 1e-3 after 300 steps, drag off."""
    d = {k: np.zeros(1, f32) for k in ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az")}
    d["x"][0] = 1.0
    d["vz"][0] = 1.0
    m = np.full(1, 1e-9, f32)
    p = O.step_params(dt=0.01, theta=0.5, softening=1e-3, collisions=0)
    s = O.ref_steps(d, m, p, 0, 300, init_forces=True)
    r = math.sqrt(float(s["x"][0]) ** 2 + float(s["y"][0]) ** 2 + float(s["z"][0]) ** 2)
    e = 0.5 * (float(s["vx"][0]) ** 2 + float(s["vy"][0]) ** 2 + float(s["vz"][0]) ** 2) - 1 / r
    assert abs(e + 0.5) < 1e-3


def test_mt_force_target_subset_equals_full_rows():
    """The walk over a target list gives exactly the rows of the full walk
    (same tree, same per-body walk), counts included."""
    from helpers import disk
    d = disk(6000, seed=17)
    full = O.mt_force(d["x"], d["y"], d["z"], d["mass"], 0.5, 1e-3, counts=True)
    tg = np.random.default_rng(3).choice(6000, 700, replace=False)
    sub = O.mt_force(d["x"], d["y"], d["z"], d["mass"], 0.5, 1e-3, counts=True, targets=tg)
    np.testing.assert_array_equal(sub["self_a"], full["self_a"][tg])
    np.testing.assert_array_equal(sub["n_acc"], full["n_acc"][tg])
    np.testing.assert_array_equal(sub["n_open"], full["n_open"][tg])


def _planet_formation_frames(seed, frames=3, exact=1):
    """The reference's planet_formation.json run (N = 10,000, r_coll = 500
    physical radii = 0.0261 AU, restitution 0.05, drag 0.01) through the
    oracle's restated step loop from this library's disk generator:
    (n_agg, largest aggregate) at the first `frames` output frames (every
    10 steps), FoF at link length r_coll as main.cu:148."""
    import nbx
    c = nbx.loadConfig(os.path.join(GOLDEN, "config_planet_formation.json"))
    r_coll = c.collision_radius_factor * nbx.config_particle_radius(c)
    d = nbx.initDisk(c, seed)
    p = O.step_params(dt=c.dt, theta=c.theta, softening=c.softening_AU, star_mass=c.star_mass_msun,
                      r_coll=r_coll, restitution=c.restitution, gas_drag_rate=c.gas_drag_rate,
                      perturber_enabled=0, collisions=1, exact_collide=exact)
    s = O.ref_steps(d, d["mass"], p, 0, 0, init_forces=True)
    out = []
    for fr in range(frames):
        s = O.ref_steps(s, d["mass"], p, 10 * fr, 10)
        lab, _ = O.fof(s["x"], s["y"], s["z"], r_coll)
        b = np.bincount(lab)
        sizes = b[b >= 2]
        out.append((len(sizes), int(sizes.max())))
    return out


def reference_control_frames():
    """First rows of the reference's own analysis of its planet_formation run
    (analysis_control/per_frame_summary.csv): (n_agg, max_mass) per frame."""
    rows = open(os.path.join(GOLDEN, "control_per_frame_summary_head.csv")).read().split()[1:]
    return [(int(r.split(",")[3]), int(r.split(",")[4])) for r in rows]


def test_planet_formation_statistics_match_reference_run():
    """Statistical sanity against the reference's only recorded output (its
    control run's first frames: 880 / 845 / 835 aggregates, largest 26 / 51
    / 33 particles).  The reference drew its disk from std::mt19937, this
    library from its own generator with the same law, so only statistics can
    agree: for three seeds every frame's aggregate count lies within 10 % of
    the reference's and the largest aggregate within a factor 3; the mean
    count over seeds and frames within 5 %."""
    ref = reference_control_frames()[:3]
    counts = []
    for seed in (42, 1, 2):
        got = _planet_formation_frames(seed)
        for (n, mx), (rn, rmx) in zip(got, ref):
            assert abs(n - rn) <= 0.10 * rn, (seed, got, ref)
            assert rmx / 3 <= mx <= 3 * rmx, (seed, got, ref)
            counts.append(n)
    mean_ref = np.mean([r[0] for r in ref])
    assert abs(np.mean(counts) - mean_ref) <= 0.05 * mean_ref, (np.mean(counts), mean_ref)
