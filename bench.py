"""Benchmark: plain Barnes-Hut disk step at N = 131,072 on one MI355X.

Workload (the repository-scale plain full step): a seeded disk of 131,072
equal-mass bodies (total 3e-5 M_sun, r in [0.5, 4] AU), theta = 0.5
(half-width / COM-distance rule), softening 1e-3 AU, dt = 0.01 T0, octree
rebuilt every step, two-pass collision impulse (restitution 0,
r_coll = particlePhysicalRadius), gas drag 1e-3, a 10 M_Jup perturber on a
circular orbit at 5.2 AU, FoF labels every 100 steps (link length r_coll).

A step is one pass of the hot path over the resident state: kick + drift,
bbox + root cube, keys, radix sort, hierarchy, node mass/COM, force walk +
perturber, collision impulse, kick + drag (plus FoF on every 100th step).
Prints ONE JSON line.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

_REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(_REPO, "nbody-sim_amd"))

METRIC = ("particle-steps/s at N=131k θ=0.5; batched sweep vs plain, "
          "rel force err≤1e-3")
N_BODIES = 131072
DISK_MASS = 3e-5
FOF_EVERY = 100
MI355X_HBM_GBPS = 8000.0
MI355X_FP32_TFLOPS = 157.3
TRAFFIC_FILE = os.path.join(_REPO, "profiles", "traffic_c2.txt")


def measured_traffic(kernel):
    """HBM bytes per dispatch of `kernel` (a name up to a template argument,
    e.g. "k_walk<32,false") from the committed counter summary (rocprofv3
    --pmc FETCH_SIZE and WRITE_SIZE, each scaled by what the same counter
    reads for a plain copy of known size); None when not measured."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                w = line.split()
                if (len(w) >= 6 and w[0] == "kernel" and w[1].startswith(kernel)
                        and w[1][len(kernel):len(kernel) + 1] in (",", ">", "")):
                    return float(w[3]) + float(w[5])
    except OSError:
        pass
    return None


BUILD_KERNELS = ["k_keys_scatter", "k_bucket_sort", "k_prefix_karras<0>", "k_oct_count",
                 "k_oct_fill"]


def build_traffic_model(n, n_rec):
    """Compulsory HBM bytes of one tree build on the step's path (key pass
    with bucket scatter, bucket sort, moment prefix and hierarchy in one
    launch, child count, record fill): every array read once by its consumer
    and written once by its producer.  n_rec: walk records (n leaves plus
    the octree nodes)."""
    n_oct = n_rec - n
    return {
        "keys: read x, y, z, m": 16 * n,
        "keys: write key, index, staged body": 28 * n,
        "bucket sort: read staged key, index, body": 28 * n,
        "bucket sort: write sorted key, index, body": 28 * n,
        "prefix: read sorted bodies": 16 * n,
        "prefix: write FP64 moments": 32 * n,
        "hierarchy: read keys": 8 * n,
        "hierarchy: write children, prefix length, flag, cell, range": 40 * n,
        "count: read children, flags": 12 * n,
        "count: write counts, first child, child codes": 4 * n + 36 * n_oct,
        "fill: read counts, first, codes, cell, range and moments of nodes, leaf bodies":
            4 * n + 132 * n_oct + 16 * n,
        "fill: write records and octree parents": 52 * n_rec,
    }


def progress(msg):
    """One progress line on stderr per finished leg (stdout holds only the
    JSON result)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--n", type=int, default=N_BODIES)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-sweep leg")
    ap.add_argument("--no-accurate", action="store_true",
                    help="skip the theta 0.25 quadrupole leg")
    ap.add_argument("--batch-steps", type=int, default=300)
    ap.add_argument("--plain-steps", type=int, default=40,
                    help="steps per disk of the reference-call-sequence sweep runs")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch the step's kernels directly instead of replaying a hipGraph")
    ap.add_argument("--no-window", action="store_true",
                    help="skip the differentiable-window leg")
    ap.add_argument("--window-steps", type=int, default=200)
    ap.add_argument("--no-precision", action="store_true",
                    help="skip the 1M-body precision study")
    ap.add_argument("--precision-steps", type=int, default=500)
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MIB",
                    help="first run a plain device copy of MIB MiB (calibrates the HBM "
                         "byte counters of a rocprofv3 --pmc run)")
    return ap.parse_args()


def host_info():
    """The timing host: CPU model, core counts, the threads this process may
    use (its affinity set capped by OMP_NUM_THREADS, the host share of one
    GPU), and the oracle's compiler and flags."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    nproc = os.cpu_count() or 1
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = nproc
    share = affinity
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        share = max(1, min(share, int(os.environ["OMP_NUM_THREADS"])))
    try:
        cc = subprocess.run(["gcc", "--version"], capture_output=True, text=True,
                            timeout=10).stdout.splitlines()[0]
    except (OSError, IndexError, subprocess.SubprocessError):
        cc = "gcc (version unknown)"
    flags = ""
    try:
        with open(os.path.join(_REPO, "oracle", "Makefile")) as f:
            for line in f:
                if line.startswith("CFLAGS"):
                    flags = line.split(":=", 1)[1].strip()
    except OSError:
        pass
    return {"cpu_model": model, "nproc": nproc, "affinity": affinity, "threads": share,
            "compiler": cc, "flags": flags}


def _median_rate(fn, units, reps=3):
    """Median of `reps` timings of fn() (after the caller's warm-up), as
    units per second, and the individual seconds."""
    secs = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        secs.append(time.perf_counter() - t0)
    med = sorted(secs)[len(secs) // 2]
    return units / med, secs


def cpu_step_rates(O, d, params, threads, target_s=3.0, reps=3):
    """Particle-steps/s of the oracle's restated reference loop
    (main.cu:101-140: canonical octree, per-body walk, collisions, drag,
    perturber) on `threads` host threads: one warm-up step, then the median
    of `reps` runs of k steps, k chosen so one run lasts about target_s."""
    n = len(d["mass"])
    s = O.ref_steps(d, d["mass"], params, 0, 0, init_forces=True, nthreads=threads)
    t0 = time.perf_counter()
    s = O.ref_steps(s, d["mass"], params, 0, 1, nthreads=threads)
    one = time.perf_counter() - t0
    k = max(1, min(100, int(target_s / max(one, 1e-4))))
    rate, secs = _median_rate(lambda: O.ref_steps(s, d["mass"], params, 1, k, nthreads=threads),
                              n * k, reps)
    return {"value": rate, "threads": threads, "steps_per_run": k, "runs_s": secs}


def cpu_baseline(n, r_coll, pert_mass_mjup, make_disk, np):
    """CPU paths of the reference timed on this host (B2 / B3 of the plan):
    the restated reference step loop at N = 131,072 and N = 4,096 on every
    host thread of this process and on one, and the restated reference
    known-answer tests.  Returns (headline, details)."""
    sys.path.insert(0, os.path.join(_REPO, "oracle"))
    import pyoracle as O
    host = host_info()
    thr = host["threads"]
    params = O.step_params(dt=0.01, theta=0.5, softening=1e-3, star_mass=1.0, r_coll=r_coll,
                           restitution=0.0, gas_drag_rate=1e-3, perturber_enabled=1,
                           pert_mass_mjup=pert_mass_mjup, pert_radius_au=5.2, collisions=1,
                           exact_collide=0)
    d = make_disk(n, DISK_MASS, 0.5, 4.0, seed=2024)
    c2_all = cpu_step_rates(O, d, params, thr, target_s=3.0)
    c2_one = cpu_step_rates(O, d, params, 1, target_s=0.6)
    m1 = np.float32(1e-5 / 4096)
    p1 = O.step_params(dt=0.01, theta=0.5, softening=1e-3, star_mass=1.0,
                       r_coll=O.particle_physical_radius(m1), restitution=0.0,
                       gas_drag_rate=1e-3, collisions=1, exact_collide=0)
    d1 = make_disk(4096, 1e-5, 0.5, 4.0, seed=2025)
    c1_all = cpu_step_rates(O, d1, p1, thr, target_s=1.0)
    c1_one = cpu_step_rates(O, d1, p1, 1, target_s=1.0)
    # the restated known-answer tests of the reference (tests/test_bhtree.cu,
    # test_collision.cu, test_integrator.cu, test_perturber.cu,
    # test_aggregate.cpp, test_disk_init.cpp), run on the CPU oracle
    t0 = time.perf_counter()
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider",
                        os.path.join(_REPO, "tests", "test_oracle.py"), "-k",
                        "kat or permutation"], capture_output=True, text=True, timeout=300,
                       cwd=_REPO)
    kat_s = time.perf_counter() - t0
    last = (r.stdout.strip().splitlines() or [""])[-1]
    headline = {"value": c2_all["value"], "unit": "particle-steps/s", "cores": thr,
                "kind": "port",
                "sample": f"median of 3 runs of {c2_all['steps_per_run']} steps of the oracle's "
                          f"restated reference loop (canonical octree, per-body walk, "
                          f"collisions, drag, perturber) at N={n} on {thr} threads "
                          f"({host['cpu_model']}, nproc {host['nproc']})"}
    details = {"host": host,
               "B2_step_N131072": {"all_threads": c2_all, "one_thread": c2_one},
               "B2_step_N4096": {"all_threads": c1_all, "one_thread": c1_one},
               "B3_reference_kats": {"ok": r.returncode == 0, "summary": last,
                                     "wall_s": kat_s}}
    return headline, details


def fof_timing(nbx, np, torch, make_disk):
    """FoF (Aggregate.cpp:26-76) on the CPU against the GPU partition of the
    same positions.  CPU: the oracle's restatement of the reference's O(N^2)
    union-find (path halving, union by rank, the float predicate), one host
    thread.  GPU: one synchronous nbx_fof call (tree build + linking +
    labels).  N = 4,096 (median of 3) and 131,072 (one call), each at link
    length r_coll and at the planet-formation 0.0261 AU, where groups are
    non-trivial."""
    sys.path.insert(0, os.path.join(_REPO, "oracle"))
    import pyoracle as O
    out = {"cores": 1, "kind": "port", "cpu": "oracle orc_fof_n2 (restated Aggregate.cpp:26-76)"}
    for n, mass in ((4096, 1e-5), (131072, DISK_MASS)):
        d = make_disk(n, mass, 0.5, 4.0, seed=2024)
        p = nbx.to_gpu(d)
        t = nbx.allocateOctree(n)
        lab = torch.empty(n, dtype=torch.int32, device="cuda")
        r_coll = nbx.particlePhysicalRadius(np.float32(mass / n))
        for ll_name, ll in (("r_coll", r_coll), ("0.0261", 0.0261)):
            if n > 4096 and ll_name == "r_coll":
                continue  # the same O(N^2) pair loop as at 0.0261 AU: one 131k call is enough
            reps = 3 if n <= 4096 else 1
            secs = []
            for _ in range(reps):
                t0 = time.perf_counter()
                labels, groups = O.fof(d["x"], d["y"], d["z"], ll, method="n2")
                secs.append(time.perf_counter() - t0)
            cpu_s = sorted(secs)[len(secs) // 2]
            nbx.fof_into(t, p, ll, lab)  # warm-up
            torch.cuda.synchronize()
            k = 20
            t0 = time.perf_counter()
            for _ in range(k):
                nbx.fof_into(t, p, ll, lab)
            torch.cuda.synchronize()
            gpu_s = (time.perf_counter() - t0) / k
            same = bool(np.array_equal(lab.cpu().numpy(), labels))
            sizes = np.bincount(labels)
            out[f"N{n}_ll_{ll_name}"] = {
                "ll": float(ll), "cpu_ms": cpu_s * 1e3, "cpu_runs": reps,
                "gpu_ms": gpu_s * 1e3, "pair_tests": n * (n - 1) // 2,
                "cpu_pair_tests_per_s": n * (n - 1) / 2 / cpu_s,
                "groups_ge2": int((sizes >= 2).sum()), "largest": int(sizes.max()),
                "labels_equal": same}
        t.free()
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU layout is described, not run", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    import nbx
    from nbx.disk import make_disk

    if not torch.cuda.is_available():
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)
    if args.copy_calib > 0:
        src = torch.ones(args.copy_calib << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(2):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    n = args.n
    d = make_disk(n, DISK_MASS, 0.5, 4.0, seed=2024)
    r_coll = nbx.particlePhysicalRadius(np.float32(DISK_MASS / n))
    sp = nbx.step_params(dt=0.01, theta=0.5, softening=1e-3, star_mass=1.0, restitution=0.0,
                         gas_drag_rate=1e-3, collisions=True, use_graph=not args.no_graph)
    pert = nbx.makePerturber(True, 10.0, 5.2, 0.0, 1.0)
    dp = nbx.disk_params(r_coll, pert)
    p = nbx.to_gpu(d)
    t = nbx.allocateOctree(n)
    labels = torch.empty(n, dtype=torch.int32, device="cuda")
    nbx.init_forces(t, p, sp, dp)
    tune_ms = nbx.tune_walk(t, 0.5, 1e-3, 1.0)
    nbx.set_async(t, True)

    nbx.step(t, p, sp, dp, 0, args.warmup)
    nbx.sync(t)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    nbx.step(t, p, sp, dp, args.warmup, args.steps)
    nbx.sync(t)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    progress(f"timed loop: {args.steps} steps in {elapsed:.3f} s")
    # FoF at every 100th step (main.cu:143-148 with output every 100 steps):
    # timed on its own and amortised into every step, so a window of any
    # length carries one FoF per 100 steps
    nbx.fof_into(t, p, r_coll, labels)
    torch.cuda.synchronize()
    k_fof = 20
    t0 = time.perf_counter()
    for _ in range(k_fof):
        nbx.fof_into(t, p, r_coll, labels)
    torch.cuda.synchronize()
    fof_ms = (time.perf_counter() - t0) / k_fof * 1e3
    loop_ms = elapsed / args.steps * 1e3
    ms_per_step = loop_ms + fof_ms / FOF_EVERY
    value = n / (ms_per_step * 1e-3)
    # the next steps rebuild their splitters from the FoF build: resettle
    nbx.step(t, p, sp, dp, args.warmup + args.steps, 5)
    nbx.sync(t)

    # per-phase kernel time with HIP events on the tree's stream (separate
    # short pass so the timed loop above stays free of event bookkeeping)
    nbx.set_timing(t, True)
    k_t = min(args.steps, 100)
    nbx.step(t, p, sp, dp, args.warmup + args.steps + 5, k_t)
    nbx.sync(t)
    phases = {ph: nbx.phase_ms(t, ph) for ph in ("kdk1", "build", "walk", "collide", "kdk2")}
    nbx.set_timing(t, False)
    n_rec = t.stats()[0][0]  # walk records of the step's tree
    # the walk's launch time on the step's tree: three back-to-back launches
    # of the step's walk kernel between HIP events on the tree's stream (the
    # per-phase events above also hold the launch gaps); the shape is kept
    shape = nbx.walk_shape(t)
    walk_ms = nbx.tune_walk(t, 0.5, 1e-3, 1.0)[shape]
    nbx.set_walk_shape(t, shape)

    # algorithmic work of the walk: 20 flops per accepted interaction and 10
    # per opened node, with the exact per-body counts of the same walk
    acc_stats = force_accuracy(nbx, t, p, n)
    walk_flops = (20.0 * acc_stats["I_acc"] + 10.0 * acc_stats["I_open"]) * n
    walk_s = walk_ms * 1e-3
    walk_tflops = walk_flops / walk_s / 1e12 if walk_s > 0 else 0.0
    # the build phase: compulsory bytes of the implemented pipeline
    build_items = build_traffic_model(n, n_rec)
    build_bytes = float(sum(build_items.values()))
    build_gbps = build_bytes / (phases["build"] * 1e-3) / 1e9 if phases["build"] > 0 else 0.0
    build_counted = [measured_traffic(k) for k in BUILD_KERNELS]
    # the walk is the step's dominant kernel; it computes on the FP32 VALU (no
    # contraction to put on the matrix cores), priced against the FP32 peak
    roof_walk = {"bound": "valu", "achieved": walk_tflops, "peak": MI355X_FP32_TFLOPS,
                 "unit": "TFLOP/s", "frac": walk_tflops / MI355X_FP32_TFLOPS,
                 "traffic": measured_traffic(f"k_walk<{shape},false"),
                 "kernel": f"k_walk<{shape},false,2,false>", "ms": walk_ms}
    roof_build = {"bound": "hbm", "achieved": build_gbps, "peak": MI355X_HBM_GBPS,
                  "unit": "GB/s", "frac": build_gbps / MI355X_HBM_GBPS,
                  "traffic": (sum(build_counted) if all(v is not None for v in build_counted)
                              else None),
                  "algorithmic_bytes": build_bytes, "records": n_rec,
                  "kernel": "build phase (" + ", ".join(BUILD_KERNELS) + ")",
                  "ms": phases["build"]}
    # achievable rates on this GPU: FMA chains and the walk's own record mix
    # (same 20-flop pricing), and a device-to-device copy
    fma_tf = nbx.probe_fp32(0)
    mix_tf = nbx.probe_fp32(1)
    src = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    copy_gbps = 5 * 2 * src.numel() * 4 / (time.perf_counter() - t0) / 1e9
    del src, dst
    achievable = {"fp32_fma_chain_tflops": fma_tf, "walk_record_mix_tflops": mix_tf,
                  "hbm_copy_gbps": copy_gbps,
                  "walk_frac_of_record_mix": walk_tflops / mix_tf if mix_tf > 0 else None,
                  "build_frac_of_copy": build_gbps / copy_gbps if copy_gbps > 0 else None}
    out = {
        "metric": METRIC, "value": value, "unit": "particle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "loop_ms_per_step": loop_ms, "fof_ms": fof_ms,
        "fof_amortised": f"one FoF (link length r_coll) per {FOF_EVERY} steps: "
                         f"ms_per_step = loop ms/step + fof_ms / {FOF_EVERY}",
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "plain full step, seeded disk N=131072 (3e-5 M_sun, "
                               "r 0.5-4 AU), theta 0.5, eps 1e-3, dt 0.01, collisions, "
                               "drag 1e-3, perturber 10 M_Jup @ 5.2 AU, FoF every 100",
                   "n": n, "theta": 0.5, "parallelism": "single GPU",
                   "launch": "direct" if args.no_graph else "hipGraph replay per step"},
        "roofline": roof_walk,
        "roofline_build": roof_build,
        "achievable": achievable,
        "phases_ms": phases,
        "walk_shape": {"targets_per_wave": nbx.walk_shape(t),
                       "tuned_ms": {str(k): v for k, v in tune_ms.items()}},
        "force_err": acc_stats,
    }
    progress(f"headline done: {value / 1e6:.1f} M particle-steps/s")
    if not args.no_accurate:
        out["accurate_step"] = accurate_step(nbx, np, torch, make_disk, args, fof_ms)
        progress("accurate step done")
    if not args.no_batch:
        out["batched_sweep"] = batched_sweep(nbx, np, torch, make_disk, args)
        progress("batched sweep done")
    if not args.no_window:
        out["diff_window"] = diff_window(nbx, np, torch, make_disk, args)
        progress("differentiable window done")
    if not args.no_precision:
        out["precision_study"] = precision_study(nbx, np, torch, make_disk, args)
        progress("precision study done")
    if not args.no_cpu:
        out["cpu_fof"] = fof_timing(nbx, np, torch, make_disk)
        progress("CPU FoF done")
        out["cpu_baseline"], out["cpu_legs"] = cpu_baseline(n, r_coll, 10.0, make_disk, np)
        progress("CPU baseline done")
    print(json.dumps(out))


SWEEP_N = [16384, 32768, 65536, 98304, 131072, 262144]
SWEEP_MJUP = [0.0, 0.3, 1.0, 3.0, 5.0, 10.0]


def plain_driver_runs(nbx, steps):
    """The six sweep disks run one after the other by the C++ driver with the
    reference's own call sequence (main.cu:85-138, one synchronous library
    call per kernel): wall time of each step loop, from the driver."""
    runs = []
    with tempfile.TemporaryDirectory() as tmp:
        for i, (n, mj) in enumerate(zip(SWEEP_N, SWEEP_MJUP)):
            cfg = {"n_particles": n, "disk_r_min": 0.5, "disk_r_max": 4.0, "disk_h_factor": 0.01,
                   "particle_mass_msun": DISK_MASS / n, "star_mass_msun": 1.0,
                   "softening_AU": 1e-3, "theta": 0.5, "collision_radius_factor": 1.0,
                   "restitution": 0.0, "gas_drag_rate": 1e-3, "dt": 0.01,
                   "total_time": 1000.0, "output_every_n_steps": 1000000, "output_dir": tmp,
                   "perturber_enabled": mj > 0, "perturber_mass_mjup": mj,
                   "perturber_radius_AU": 5.2, "perturber_phase0": 0.0, "disk_seed": 100 + i}
            path = os.path.join(tmp, f"disk{i}.json")
            with open(path, "w") as f:
                json.dump(cfg, f)
            r = subprocess.run([nbx.RUN_PATH, path, "--plain", "--no-output", "--steps",
                                str(steps)], capture_output=True, text=True, timeout=600)
            if r.returncode != 0:
                raise RuntimeError("plain driver run failed:\n" + r.stdout + r.stderr)
            runs.append(json.loads(r.stdout.strip().splitlines()[-1]))
    return runs


def batched_sweep(nbx, np, torch, make_disk, args):
    """Six disks of uneven N, one perturber mass each, advanced together:
    two batched trees over contiguous disk groups (every phase one launch
    per group), each replaying its captured step on its own stream, side by
    side, with FoF of every disk (at its own r_coll) after every 100th step
    inside the timed loop.  Against the same six disks run one after the
    other as captured single-disk steps of this library with the same FoF
    (the primary comparison), and by the C++ driver with the reference's
    own call sequence (steps only); the six disks in one batched tree on one
    stream are reported beside."""
    k = max(FOF_EVERY, -(-args.batch_steps // FOF_EVERY) * FOF_EVERY)  # whole FoF periods
    disks = [make_disk(n, DISK_MASS, 0.5, 4.0, seed=100 + i) for i, n in enumerate(SWEEP_N)]
    sp = nbx.step_params(dt=0.01, theta=0.5, softening=1e-3, star_mass=1.0, restitution=0.0,
                         gas_drag_rate=1e-3, collisions=True, use_graph=not args.no_graph)
    dps, lls = [], []
    for n, mj in zip(SWEEP_N, SWEEP_MJUP):
        r_coll = nbx.particlePhysicalRadius(np.float32(DISK_MASS / n))
        lls.append(r_coll)
        dps.append(nbx.disk_params(r_coll, nbx.makePerturber(mj > 0, mj, 5.2, 0.0, 1.0)))

    def loop(trees, parts, dlist, llist, labs, step0, nsteps, together):
        """nsteps steps from step0 with FoF of every tree after each 100th
        step (trees side by side when `together`, else one after the other)."""
        s = step0
        while s < step0 + nsteps:
            nxt = min(step0 + nsteps, (s // FOF_EVERY + 1) * FOF_EVERY)
            if together:
                nbx.step_concurrent(trees, parts, sp, dlist, s, nxt - s)
            else:
                for t, p, d in zip(trees, parts, dlist):
                    nbx.step(t, p, sp, d, s, nxt - s)
            s = nxt
            if s % FOF_EVERY == 0:
                for t, p, ll, lab in zip(trees, parts, llist, labs):
                    nbx.fof_into(t, p, ll, lab)
        for t in trees:
            nbx.sync(t)
        torch.cuda.synchronize()

    # each disk alone: its own tree and captured step, one after the other
    plain = []
    for dd, dp in zip(disks, dps):
        p = nbx.to_gpu(dd)
        t = nbx.allocateOctree(p.n)
        nbx.init_forces(t, p, sp, dp)
        nbx.tune_walk(t, 0.5, 1e-3, 1.0)
        nbx.set_async(t, True)
        plain.append((t, p, dp))
    pt, pp, pd = ([m[j] for m in plain] for j in range(3))
    plabs = [torch.empty(p.n, dtype=torch.int32, device="cuda") for p in pp]
    loop(pt, pp, pd, lls, plabs, 0, 5, False)  # warm-up (graph capture)
    t0 = time.perf_counter()
    loop(pt, pp, pd, lls, plabs, 5, k, False)
    seq_s = time.perf_counter() - t0
    # steps only (for the comparison with the reference call sequence)
    t0 = time.perf_counter()
    for t, p, d in plain:
        nbx.step(t, p, sp, d, 5 + k, FOF_EVERY)
        nbx.sync(t)
    torch.cuda.synchronize()
    seq_steps_s = (time.perf_counter() - t0) / FOF_EVERY
    # the six disks in two contiguous groups (even body totals), each one
    # batched tree replaying its captured step on its own stream, side by side
    groups = nbx.split_groups(SWEEP_N, 2)
    members = []
    for g in groups:
        cat = {key: np.concatenate([disks[i][key] for i in g]) for key in disks[0]}
        pg = nbx.to_gpu(cat)
        tg = nbx.allocateOctreeBatch([SWEEP_N[i] for i in g])
        dg = [dps[i] for i in g]
        nbx.init_forces(tg, pg, sp, dg)
        nbx.tune_walk(tg, 0.5, 1e-3, 1.0)
        nbx.set_async(tg, True)
        members.append((tg, pg, dg, [lls[i] for i in g]))
    mt, mp, md, ml = ([m[j] for m in members] for j in range(4))
    mlabs = [torch.empty(p.n, dtype=torch.int32, device="cuda") for p in mp]
    loop(mt, mp, md, ml, mlabs, 0, 5, True)
    t0 = time.perf_counter()
    loop(mt, mp, md, ml, mlabs, 5, k, True)
    bat_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    nbx.step_concurrent(mt, mp, sp, md, 5 + k, FOF_EVERY)
    for tg in mt:
        nbx.sync(tg)
    torch.cuda.synchronize()
    bat_steps_s = (time.perf_counter() - t0) / FOF_EVERY
    # all six disks in one batched tree (one stream)
    cat = {key: np.concatenate([dd[key] for dd in disks]) for key in disks[0]}
    pb = nbx.to_gpu(cat)
    tb = nbx.allocateOctreeBatch(SWEEP_N)
    nbx.init_forces(tb, pb, sp, dps)
    nbx.tune_walk(tb, 0.5, 1e-3, 1.0)
    nbx.set_async(tb, True)
    blab = [torch.empty(pb.n, dtype=torch.int32, device="cuda")]
    loop([tb], [pb], [dps], [lls], blab, 0, 5, False)
    t0 = time.perf_counter()
    loop([tb], [pb], [dps], [lls], blab, 5, k, False)
    one_s = time.perf_counter() - t0
    total = float(sum(SWEEP_N))
    batched = total * k / bat_s
    runs = plain_driver_runs(nbx, args.plain_steps)
    plain_rate = sum(r["n"] * r["steps"] for r in runs) / sum(r["loop_s"] for r in runs)
    return {"disks": SWEEP_N, "perturber_mjup": SWEEP_MJUP, "steps": k,
            "fof": f"every disk at its own r_coll after every {FOF_EVERY}th step, in the loop",
            "groups": groups,
            "how": "two batched trees (contiguous disk groups), each replaying its "
                   "captured step on its own stream, side by side (nbx_step_concurrent)",
            "walk_shape_batched": [nbx.walk_shape(tg) for tg in mt],
            "walk_shape_captured_runs": [nbx.walk_shape(t) for t in pt],
            "batched_particle_steps_per_s": batched,
            "batched_ms_per_step": bat_s / k * 1e3,
            "speedup_vs_captured_runs": seq_s / bat_s,
            "captured_runs": {"how": "each disk alone, one replayed step graph, the same FoF",
                              "particle_steps_per_s": total * k / seq_s,
                              "ms_per_step": seq_s / k * 1e3},
            "single_tree": {"how": "all six disks in one batched tree on one stream",
                            "particle_steps_per_s": total * k / one_s,
                            "ms_per_step": one_s / k * 1e3},
            "steps_only": {"batched_ms_per_step": bat_steps_s * 1e3,
                           "captured_runs_ms_per_step": seq_steps_s * 1e3,
                           "speedup_vs_captured_runs": seq_steps_s / bat_steps_s,
                           "speedup_vs_plain_runs": total / bat_steps_s / plain_rate},
            "plain_runs": {"how": "C++ driver, reference call sequence main.cu:85-138, "
                                  "one synchronous call per kernel, steps only",
                           "steps_per_disk": args.plain_steps,
                           "particle_steps_per_s": plain_rate,
                           "ms_per_step_per_disk": [r["loop_s"] / r["steps"] * 1e3
                                                    for r in runs]}}


def accurate_step(nbx, np, torch, make_disk, args, fof_ms):
    """The plain full step at the accuracy the metric names: the same
    131,072-body workload with theta 0.25 and the quadrupole of every
    accepted node (median self-gravity error below 1e-3 against the FP64
    direct sum; the reference's theta 0.5 monopole rule is several percent
    off).  Throughput with FoF amortised as in the headline, the walk's
    time and its FP32 work priced like the headline walk (20 flop per
    accepted and 10 per opened record at theta 0.25; the quadrupole terms
    are not counted), and the force error."""
    n = args.n
    d = make_disk(n, DISK_MASS, 0.5, 4.0, seed=2024)
    r_coll = nbx.particlePhysicalRadius(np.float32(DISK_MASS / n))
    theta = 0.25
    sp = nbx.step_params(dt=0.01, theta=theta, softening=1e-3, star_mass=1.0, restitution=0.0,
                         gas_drag_rate=1e-3, collisions=True, use_graph=not args.no_graph)
    dp = nbx.disk_params(r_coll, nbx.makePerturber(True, 10.0, 5.2, 0.0, 1.0))
    p = nbx.to_gpu(d)
    t = nbx.allocateOctree(n)
    nbx.set_multipole(t, 2)
    nbx.init_forces(t, p, sp, dp)
    tune = nbx.tune_walk(t, theta, 1e-3, 1.0)
    nbx.set_async(t, True)
    nbx.step(t, p, sp, dp, 0, args.warmup)
    nbx.sync(t)
    torch.cuda.synchronize()
    k = max(args.steps, 20)
    t0 = time.perf_counter()
    nbx.step(t, p, sp, dp, args.warmup, k)
    nbx.sync(t)
    torch.cuda.synchronize()
    loop_ms = (time.perf_counter() - t0) / k * 1e3
    nbx.set_timing(t, True)
    nbx.step(t, p, sp, dp, args.warmup + k, 20)
    nbx.sync(t)
    phases = {ph: nbx.phase_ms(t, ph) for ph in ("kdk1", "build", "walk", "collide", "kdk2")}
    nbx.set_timing(t, False)
    shape = nbx.walk_shape(t)
    walk_ms = nbx.tune_walk(t, theta, 1e-3, 1.0)[shape]  # back-to-back launches, as the headline
    nbx.set_walk_shape(t, shape)
    acc, opn = nbx.walk_counts(t, p, theta)
    flops = (20.0 * acc + 10.0 * opn) * n
    tf = flops / (walk_ms * 1e-3) / 1e12
    tg = torch.from_numpy(np.sort(np.random.default_rng(0).choice(n, 1024, replace=False)))
    tg = tg.cuda()
    err = walk_errors(nbx, t, p, n, theta, tg, fp64_reference(p, tg), quad=True)
    ms = loop_ms + fof_ms / FOF_EVERY
    return {"theta": theta, "multipole": 2, "particle_steps_per_s": n / (ms * 1e-3),
            "ms_per_step": ms, "loop_ms_per_step": loop_ms, "steps": k, "phases_ms": phases,
            "walk_shape": nbx.walk_shape(t), "tuned_ms": {str(a): b for a, b in tune.items()},
            "I_acc": acc, "I_open": opn,
            "roofline": {"bound": "valu", "achieved": tf, "peak": MI355X_FP32_TFLOPS,
                         "unit": "TFLOP/s", "frac": tf / MI355X_FP32_TFLOPS,
                         "kernel": f"k_walk<{shape},true,1,false>", "ms": walk_ms},
            "force_err": err}


def diff_window(nbx, np, torch, make_disk, args):
    """Differentiable window: seeded disk of 32,768 bodies (3e-5 M_sun), no
    collisions, gravity + drag 1e-3 + 10 M_Jup perturber at 5.2 AU, 200
    steps with a checkpoint every 20; loss = mean e^2 + i^2 of the bodies in
    1-2 AU at the end.  Times the forward window and forward + reverse, and
    checks d loss / d(perturber mass, radius, v0 along a random direction)
    against central differences (relative step 1e-4) of the same window run
    through the FP64 step."""
    from nbx.window import (WindowConfig, annulus_mask, run_window, run_window64, window,
                            window_loss)
    n = 32768
    d = make_disk(n, DISK_MASS, 0.5, 4.0, seed=77)
    x0 = torch.tensor(np.stack([d["x"], d["y"], d["z"]]), device="cuda").contiguous()
    v0 = torch.tensor(np.stack([d["vx"], d["vy"], d["vz"]]), device="cuda").contiguous()
    m = torch.tensor(d["mass"], device="cuda")
    t = nbx.allocateOctree(n)
    cfg = WindowConfig(n_steps=args.window_steps, checkpoint_every=20)
    run_window(x0, v0, 10.0, 5.2, m, t, cfg)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    xT, vT, _ = run_window(x0, v0, 10.0, 5.2, m, t, cfg)
    torch.cuda.synchronize()
    fwd_ms = (time.perf_counter() - t0) * 1e3
    mask = annulus_mask(xT)

    def forward_reverse():
        pm = torch.tensor(10.0, dtype=torch.float64, requires_grad=True)
        rp = torch.tensor(5.2, dtype=torch.float64, requires_grad=True)
        v0g = v0.clone().requires_grad_(True)
        xT, vT = window(x0, v0g, pm, rp, m, t, cfg)
        L = window_loss(xT, vT, mask)
        L.backward()
        return pm, rp, v0g, L

    forward_reverse()  # warm-up: reverse-mode buffers and the first launches
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pm, rp, v0g, L = forward_reverse()
    torch.cuda.synchronize()
    fb_ms = (time.perf_counter() - t0) * 1e3

    def loss_at(pmv, rpv, vv=v0):
        # the FP64 step's rounding (~1e-16) leaves a relative step of 1e-4
        # clean; the float window's (~1e-7 per step) would not
        x, v = run_window64(x0, vv, pmv, rpv, m, t, cfg)
        return window_loss(x, v, mask).item()

    h = 1e-4
    fd_m = (loss_at(10.0 * (1 + h), 5.2) - loss_at(10.0 * (1 - h), 5.2)) / (2 * h * 10.0)
    fd_r = (loss_at(10.0, 5.2 * (1 + h)) - loss_at(10.0, 5.2 * (1 - h))) / (2 * h * 5.2)
    dv = torch.randn(3, n, generator=torch.Generator(device="cuda").manual_seed(3),
                     device="cuda")
    fd_v = (loss_at(10.0, 5.2, v0 + h * dv) - loss_at(10.0, 5.2, v0 - h * dv)) / (2 * h)
    return {"n": n, "steps": cfg.n_steps, "checkpoint_every": cfg.checkpoint_every,
            "bodies_in_loss": int(mask.sum().item()), "loss": L.item(),
            "forward_ms": fwd_ms, "forward_reverse_ms": fb_ms,
            "fd": "central, relative step 1e-4, forward through the FP64 step",
            "dL_dmass_mjup": pm.grad.item(), "fd_dmass_mjup": fd_m,
            "dL_dradius_au": rp.grad.item(), "fd_dradius_au": fd_r,
            "dL_dv0_along_random_dir": (v0g.grad.double() * dv.double()).sum().item(),
            "fd_dv0_along_random_dir": fd_v,
            "dL_dv0_norm": torch.linalg.norm(v0g.grad.double()).item()}


def fp64_direct(pos, m, tg, eps):
    """FP64 softened direct-sum self-gravity at the targets tg (torch, GPU)."""
    import torch
    chunk = max(1024, (1 << 26) // len(tg))
    acc = torch.zeros(len(tg), 3, dtype=torch.float64, device="cuda")
    pt = pos[tg]
    for c0 in range(0, pos.shape[0], chunk):
        src = pos[c0:c0 + chunk]
        dd = src[None, :, :] - pt[:, None, :]
        r2 = (dd * dd).sum(-1) + eps * eps
        inv = m[c0:c0 + chunk][None, :] / (r2 * r2.sqrt())
        same = (torch.arange(c0, c0 + src.shape[0], device="cuda")[None, :] == tg[:, None])
        acc += (dd * inv.masked_fill(same, 0.0)[..., None]).sum(1)
    return acc


def fp64_reference(p, targets, eps=1e-3):
    import torch
    pos = torch.stack([p.x, p.y, p.z], 1).double()
    return fp64_direct(pos, p.mass.double(), targets, eps)


def walk_errors(nbx, t, p, n, theta, targets, ref, quad=False, eps=1e-3):
    """Relative error of the GPU self-gravity and of the total acceleration
    (star + self-gravity) against FP64 (ref: fp64_reference) at the targets."""
    import numpy as np
    import torch
    sa = [torch.zeros(n, dtype=torch.float32, device="cuda") for _ in range(3)]
    if quad:
        nbx.set_multipole(t, 2)
    nbx.launchBBoxKernel(t, p)
    nbx.launchBuildKernel(t, p)
    nbx.launchSummarizeKernel(t)
    nbx.launchForceKernel(t, p, theta, eps, 1.0, self_acc=sa)
    if quad:
        nbx.set_multipole(t, 1)
    g = torch.stack(sa, 1).double()[targets]
    tot = torch.stack([p.ax, p.ay, p.az], 1).double()[targets]
    pt = torch.stack([p.x, p.y, p.z], 1).double()[targets]
    r2 = (pt * pt).sum(1, keepdim=True) + eps * eps
    ref_tot = ref - pt / (r2 * r2.sqrt())
    es = ((g - ref).norm(dim=1) / ref.norm(dim=1)).cpu().numpy()
    et = ((tot - ref_tot).norm(dim=1) / ref_tot.norm(dim=1)).cpu().numpy()
    q = lambda e: {"median": float(np.median(e)), "p99": float(np.percentile(e, 99))}
    return {"self_gravity": q(es), "total_acceleration": q(et)}


def force_accuracy(nbx, t, p, n, n_targets=1024):
    """Error of the GPU walk against an FP64 direct sum (torch float64 on the
    GPU) over seeded targets: the reference rule (theta 0.5, monopoles) and
    the accuracy option (theta 0.25 with quadrupoles); and the mean
    interaction counts of the theta 0.5 walk (exact per-body counts)."""
    import numpy as np
    import torch
    tg = torch.from_numpy(np.sort(np.random.default_rng(0).choice(n, n_targets, replace=False)))
    tg = tg.cuda()
    ref = fp64_reference(p, tg)
    ref_rule = walk_errors(nbx, t, p, n, 0.5, tg, ref)
    counts = nbx_counts(nbx, t, p, n)  # on the tree of the theta 0.5 walk
    accurate = walk_errors(nbx, t, p, n, 0.25, tg, ref, quad=True)
    return {"targets": n_targets, "reference_rule": "h/r_com < theta (BHTree.cu:577)",
            "theta_0.5_monopole": ref_rule, "theta_0.25_quadrupole": accurate,
            "I_acc": counts[0], "I_open": counts[1]}


def precision_study(nbx, np, torch, make_disk, args):
    """Precision study on a 1,048,576-body disk (r 0.3-10 AU): walk error
    against the FP64 direct sum for theta 0.3 / 0.5 / 0.7 on 16,384 seeded
    targets; then 500 gravity-only steps (theta 0.5, star + self-gravity) of
    the same initial state with the float step, the compensated KDK (forces
    FP32) and the all-FP64 step (the reference): each run's drift of the FP64
    energy (kinetic + star + pairwise self potential) and of L_y, and the
    float and compensated runs' position / velocity error against the FP64
    run."""
    n = 1 << 20
    d = make_disk(n, DISK_MASS, 0.3, 10.0, seed=99)
    p = nbx.to_gpu(d)
    t = nbx.allocateOctree(n)
    tg = torch.from_numpy(np.sort(np.random.default_rng(1).choice(n, 16384, replace=False)))
    tg = tg.cuda()
    out = {"n": n, "targets": 16384, "theta": {}}
    ref = fp64_reference(p, tg)
    for th in (0.3, 0.5, 0.7):
        e = walk_errors(nbx, t, p, n, th, tg, ref)
        nbx.launchBBoxKernel(t, p)
        nbx.launchBuildKernel(t, p)
        nbx.launchSummarizeKernel(t)
        e["walk_ms"] = nbx.tune_walk(t, th, 1e-3, 1.0)[nbx.walk_shape(t)]
        out["theta"][str(th)] = e
    t.free()
    sp = nbx.step_params(dt=0.01, theta=0.5, softening=1e-3, star_mass=1.0, gas_drag_rate=0.0,
                         collisions=False, use_graph=True)
    dp = nbx.disk_params(0.0, None)
    steps = args.precision_steps
    final = {}
    e_start = None
    for mode in ("fp64", "float", "compensated"):
        q = nbx.to_gpu(d)
        t2 = nbx.allocateOctree(n)
        nbx.set_precision(t2, mode)
        nbx.init_forces(t2, q, sp, dp)
        d0 = nbx.diagnostics(t2, q, with_self=e_start is None)
        if e_start is None:  # the same initial state for every mode
            w0 = d0["self_potential"]
            e_start = d0["kinetic"] + d0["star_potential"] + w0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        nbx.step(t2, q, sp, dp, 0, steps)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        d1 = nbx.diagnostics(t2, q, with_self=True)
        e1 = d1["kinetic"] + d1["star_potential"] + d1["self_potential"]
        e0_ks = d0["kinetic"] + d0["star_potential"]
        final[mode] = nbx.state64(t2, q)[:6].clone()
        out[mode] = {"steps": steps, "ms_per_step": ms,
                     "energy_rel_drift": (e1 - e_start) / abs(e_start),
                     "energy_without_self_rel_drift":
                         (d1["kinetic"] + d1["star_potential"] - e0_ks) / abs(e0_ks),
                     "self_potential_change": d1["self_potential"] - w0,
                     "Ly_rel_drift": (d1["L"][1] - d0["L"][1]) / abs(d0["L"][1])}
        t2.free()
        del q
    # the FP64 run again at dt / 2 over the same time: a drift that falls as
    # dt^2 is the integrator's truncation error, not rounding
    sp2 = nbx.step_params(dt=0.005, theta=0.5, softening=1e-3, star_mass=1.0, gas_drag_rate=0.0,
                          collisions=False, use_graph=True)
    q = nbx.to_gpu(d)
    t2 = nbx.allocateOctree(n)
    nbx.set_precision(t2, "fp64")
    nbx.init_forces(t2, q, sp2, dp)
    nbx.step(t2, q, sp2, dp, 0, 2 * steps)
    d1 = nbx.diagnostics(t2, q, with_self=True)
    out["fp64_half_dt"] = {
        "steps": 2 * steps, "dt": 0.005,
        "energy_rel_drift":
            (d1["kinetic"] + d1["star_potential"] + d1["self_potential"] - e_start) / abs(e_start)}
    t2.free()
    del q
    ref64 = final["fp64"]
    for mode in ("float", "compensated"):
        dx = (final[mode][:3] - ref64[:3]).norm(dim=0)
        dv = (final[mode][3:] - ref64[3:]).norm(dim=0)
        out[mode]["vs_fp64"] = {
            "pos_err_rms_au": float(dx.square().mean().sqrt()),
            "pos_err_max_au": float(dx.max()),
            "vel_err_rms": float(dv.square().mean().sqrt()),
            "energy_minus_fp64": out[mode]["energy_rel_drift"] - out["fp64"]["energy_rel_drift"]}
    out["energy"] = "kinetic + softened star potential + softened pairwise self potential (FP64)"
    return out


def nbx_counts(nbx, t, p, n):
    """Mean accepted / opened counts per body of the theta=0.5 walk, from the
    walk-count kernel of the library (exact per-body counts)."""
    return nbx.walk_counts(t, p, 0.5)


if __name__ == "__main__":
    main()
