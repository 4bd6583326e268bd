"""Python side of the MI355X Barnes-Hut disk step.

A thin ctypes binding of the C ABI in include/nbx.h (libnbx.so, built from
nbody-sim_amd/csrc by :func:`build`).  Device memory comes from PyTorch-ROCm
tensors; nothing here computes physics on the host.  The function names
mirror the reference's call surface (BHTree.cuh:38-50, Integrator.cuh:9-22,
Perturber.h:39-75, Aggregate.h:17-18, ParticleData.h:35-123) so code written
against the reference reads the same.

The HIP library is mandatory: if libnbx.so is missing or cannot be loaded,
every entry point raises RuntimeError (there is no CPU fallback).
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_PKG = os.path.dirname(_HERE)
_REPO = os.path.dirname(_PKG)
CSRC = os.path.join(_PKG, "csrc")
LIB_PATH = os.path.join(_HERE, "libnbx.so")
RUN_PATH = os.path.join(_HERE, "nbx_run")
HIP_SOURCES = ["nbx_tree.hip", "nbx_walk.hip", "nbx_fp64.hip", "nbx_adjoint.hip", "nbx_abi.hip",
               "nbx_io.cpp"]
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # keep hand-written float code on single-lane VALU ops (no v_pk_*_f32)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]

MJUP_TO_MSUN = 9.5479e-4  # Perturber.h:25


def build(verbose=False):
    """Compile libnbx.so for gfx950 in-tree (hipcc cross-compiles without a GPU)."""
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES]
    tmp = LIB_PATH + ".tmp"
    cmd = ["hipcc", *HIPCC_FLAGS, "-o", tmp, *srcs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB_PATH)
    # the C++ driver of src/main.cu, linked against the library beside it
    run = ["g++", "-O2", "-std=c++17", "-Wall", "-o", RUN_PATH + ".tmp",
           os.path.join(CSRC, "nbx_run.cpp"), "-L", _HERE, "-lnbx", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(run, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("driver build failed:\n" + r.stdout + r.stderr)
    os.replace(RUN_PATH + ".tmp", RUN_PATH)
    if verbose:
        print("built", LIB_PATH, "and", RUN_PATH)
    return LIB_PATH


class nbx_particles(C.Structure):
    _fields_ = [("x", C.c_void_p), ("y", C.c_void_p), ("z", C.c_void_p),
                ("vx", C.c_void_p), ("vy", C.c_void_p), ("vz", C.c_void_p),
                ("ax", C.c_void_p), ("ay", C.c_void_p), ("az", C.c_void_p),
                ("mass", C.c_void_p), ("id", C.c_void_p), ("n", C.c_int32)]


class nbx_config(C.Structure):
    """Config (Config.h:5-35) plus the generator fields disk_seed, e_sigma, i_sigma."""
    _fields_ = [("n_particles", C.c_int32), ("disk_r_min", C.c_float),
                ("disk_r_max", C.c_float), ("disk_h_factor", C.c_float),
                ("particle_mass_msun", C.c_float), ("star_mass_msun", C.c_float),
                ("softening_AU", C.c_float), ("theta", C.c_float),
                ("collision_radius_factor", C.c_float), ("restitution", C.c_float),
                ("gas_drag_rate", C.c_float), ("dt", C.c_float), ("total_time", C.c_float),
                ("output_every_n_steps", C.c_int32), ("output_dir", C.c_char * 512),
                ("perturber_enabled", C.c_int32), ("perturber_mass_mjup", C.c_float),
                ("perturber_radius_AU", C.c_float), ("perturber_phase0", C.c_float),
                ("disk_seed", C.c_uint32), ("e_sigma", C.c_float), ("i_sigma", C.c_float)]


class nbx_aggregate(C.Structure):
    _fields_ = [("id", C.c_int32), ("n_particles", C.c_int32), ("mass_msun", C.c_float),
                ("cx", C.c_float), ("cy", C.c_float), ("cz", C.c_float), ("vx", C.c_float),
                ("vy", C.c_float), ("vz", C.c_float), ("dist_center", C.c_float)]


class nbx_step_params(C.Structure):
    _fields_ = [("dt", C.c_float), ("theta", C.c_float), ("softening", C.c_float),
                ("star_mass", C.c_float), ("restitution", C.c_float),
                ("gas_drag_rate", C.c_float), ("collisions", C.c_int32),
                ("use_graph", C.c_int32)]


class nbx_disk_params(C.Structure):
    _fields_ = [("r_coll", C.c_float), ("perturber_enabled", C.c_int32),
                ("pert_mass_msun", C.c_float), ("pert_radius_au", C.c_float),
                ("pert_phase0", C.c_float), ("pert_omega", C.c_float)]


# every symbol include/nbx.h declares (checked by the CPU test suite)
EXPORTS = [
    "nbx_last_error", "nbx_particles_alloc", "nbx_particles_free", "nbx_copy_to_device",
    "nbx_copy_to_host", "nbx_tree_create", "nbx_tree_create_batch", "nbx_tree_destroy",
    "nbx_tree_reset", "nbx_tree_set_stream", "nbx_tree_set_async", "nbx_tree_sync", "nbx_bbox",
    "nbx_build", "nbx_summarize", "nbx_sort", "nbx_sort_idx", "nbx_sorted_keys", "nbx_force",
    "nbx_force_ex", "nbx_collide", "nbx_perturber_force", "nbx_half_kick", "nbx_drift",
    "nbx_reset_acceleration", "nbx_gas_drag", "nbx_fof", "nbx_init_forces", "nbx_step",
    "nbx_step_concurrent", "nbx_tree_stats", "nbx_tree_set_timing", "nbx_phase_ms", "nbx_probe_fp32", "nbx_walk_counts",
    "nbx_tree_set_multipole", "nbx_tree_set_precision", "nbx_diagnostics",
    "nbx_tree_set_walk_shape", "nbx_tree_get_walk_shape", "nbx_tune_walk", "nbx_walk_stats",
    "nbx_tree_set_radix_bits", "nbx_tree_set_adjoint", "nbx_adj_force", "nbx_adj_step",
    "nbx_adj_init_forces", "nbx_walk_wave_stats", "nbx_config_default", "nbx_config_load",
    "nbx_particle_radius", "nbx_init_disk", "nbx_writer_open", "nbx_writer_frame",
    "nbx_writer_stats", "nbx_writer_perturber", "nbx_writer_finalize", "nbx_writer_close",
    "nbx_write_perturber_sidecar", "nbx_fof_noalloc", "nbx_fof_disks", "nbx_search_stats",
    "nbx_tree_set_wave_order", "nbx_state64",
]

_lib = None


def lib():
    """Load libnbx.so (raises RuntimeError when it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP extension not built: {LIB_PATH} missing (run nbx.build())")
        L = C.CDLL(LIB_PATH)
        vp, i32, f32 = C.c_void_p, C.c_int32, C.c_float
        P = C.POINTER(nbx_particles)
        L.nbx_last_error.restype = C.c_char_p
        sig = {
            "nbx_particles_alloc": [P, i32], "nbx_particles_free": [P],
            "nbx_copy_to_device": [P, P], "nbx_copy_to_host": [P, P],
            "nbx_tree_create": [C.POINTER(vp), i32],
            "nbx_tree_create_batch": [C.POINTER(vp), i32, C.POINTER(i32)],
            "nbx_tree_destroy": [vp], "nbx_tree_reset": [vp],
            "nbx_tree_set_stream": [vp, vp], "nbx_tree_set_async": [vp, i32],
            "nbx_tree_sync": [vp], "nbx_bbox": [vp, P], "nbx_build": [vp, P],
            "nbx_summarize": [vp], "nbx_sort": [vp], "nbx_sort_idx": [vp, vp],
            "nbx_sorted_keys": [vp, vp], "nbx_force": [vp, P, f32, f32, f32],
            "nbx_force_ex": [vp, P, f32, f32, f32, vp, vp, vp],
            "nbx_collide": [vp, P, f32, f32], "nbx_perturber_force": [P, f32, f32, f32, f32, f32],
            "nbx_half_kick": [P, f32], "nbx_drift": [P, f32], "nbx_reset_acceleration": [P],
            "nbx_gas_drag": [P, f32, f32, f32], "nbx_fof": [vp, P, f32, vp],
            "nbx_fof_noalloc": [P, f32, vp, vp],
            "nbx_fof_disks": [vp, P, C.POINTER(f32), vp],
            "nbx_search_stats": [vp, C.POINTER(i32), i32],
            "nbx_tree_set_wave_order": [vp, i32],
            "nbx_init_forces": [vp, P, C.POINTER(nbx_step_params), C.POINTER(nbx_disk_params)],
            "nbx_step": [vp, P, C.POINTER(nbx_step_params), C.POINTER(nbx_disk_params), i32, i32],
            "nbx_step_concurrent": [C.POINTER(vp), C.POINTER(P), i32, C.POINTER(nbx_step_params),
                                    C.POINTER(C.POINTER(nbx_disk_params)), i32, i32],
            "nbx_tree_stats": [vp, vp, vp], "nbx_tree_set_timing": [vp, i32],
            "nbx_phase_ms": [vp, C.c_char_p, C.POINTER(f32)],
            "nbx_probe_fp32": [i32, C.POINTER(C.c_double)],
            "nbx_walk_counts": [vp, f32, vp, vp],
            "nbx_tree_set_multipole": [vp, i32],
            "nbx_tree_set_precision": [vp, i32],
            "nbx_state64": [vp, P, vp],
            "nbx_diagnostics": [vp, P, f32, f32, i32, C.POINTER(C.c_double)],
            "nbx_tree_set_walk_shape": [vp, i32],
            "nbx_tree_get_walk_shape": [vp, C.POINTER(i32)],
            "nbx_tune_walk": [vp, f32, f32, f32, C.POINTER(f32)],
            "nbx_walk_stats": [vp, f32, C.POINTER(C.c_double)],
            "nbx_tree_set_radix_bits": [vp, i32],
            "nbx_tree_set_adjoint": [vp, i32],
            "nbx_walk_wave_stats": [vp, f32, vp, i32, C.POINTER(i32)],
            "nbx_config_default": [C.POINTER(nbx_config)],
            "nbx_config_load": [C.c_char_p, C.POINTER(nbx_config)],
            "nbx_init_disk": [C.POINTER(nbx_config), C.c_uint32, P],
            "nbx_writer_open": [C.POINTER(vp), C.c_char_p, i32, f32, i32],
            "nbx_writer_frame": [vp, P, f32, vp],
            "nbx_writer_stats": [vp, i32, f32, i32, C.POINTER(nbx_aggregate)],
            "nbx_writer_perturber": [vp, i32, f32, i32, f32, f32, f32, f32, f32],
            "nbx_writer_finalize": [vp, i32],
            "nbx_writer_close": [vp],
            "nbx_write_perturber_sidecar": [C.c_char_p, i32, f32, f32, f32, f32],
            "nbx_adj_force": [vp, P, f32, f32, f32, C.POINTER(nbx_disk_params), f32, vp, vp,
                              C.POINTER(C.c_double)],
            "nbx_adj_step": [vp, P, P, C.POINTER(nbx_step_params), C.POINTER(nbx_disk_params),
                             i32, vp, vp, vp, C.POINTER(C.c_double)],
            "nbx_adj_init_forces": [vp, P, C.POINTER(nbx_step_params),
                                    C.POINTER(nbx_disk_params), vp, vp, C.POINTER(C.c_double)],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_int
        L.nbx_particle_radius.argtypes = [C.POINTER(nbx_config)]
        L.nbx_particle_radius.restype = C.c_float
        _lib = L
    return _lib


def _check(rc, what):
    if rc != 0:
        msg = lib().nbx_last_error().decode(errors="replace")
        raise RuntimeError(f"{what} failed ({rc}): {msg}")


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("no GPU visible: the nbx step runs only on the HIP device")
    return torch


# ---------------------------------------------------------------------------
# ParticleData (ParticleData.h:19-123)
# ---------------------------------------------------------------------------
FIELDS = ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az", "mass")


@dataclass
class ParticleData:
    """SoA particle state in HBM (torch tensors on the HIP device)."""
    n: int
    t: dict = field(default_factory=dict)

    def c(self):
        s = nbx_particles()
        for k in FIELDS:
            setattr(s, k, self.t[k].data_ptr())
        s.id = self.t["id"].data_ptr()
        s.n = self.n
        return s

    def __getattr__(self, k):
        if k in FIELDS or k == "id":
            return self.t[k]
        raise AttributeError(k)

    def numpy(self):
        _torch().cuda.synchronize()
        return {k: v.cpu().numpy() for k, v in self.t.items()}


def allocateParticlesGPU(n):
    torch = _torch()
    if n <= 0:
        raise RuntimeError("allocateParticlesGPU: n must be positive")
    p = ParticleData(n)
    for k in FIELDS:
        p.t[k] = torch.zeros(n, dtype=torch.float32, device="cuda")
    p.t["id"] = torch.arange(n, dtype=torch.int32, device="cuda")
    return p


def copyToGPU(cpu, gpu):
    """Host dict of numpy arrays -> device ParticleData (ParticleData.h:90-105)."""
    torch = _torch()
    n = len(cpu["x"])
    if n != gpu.n:
        raise RuntimeError(f"copyToGPU: CPU n={n} != GPU n={gpu.n}")
    for k in FIELDS:
        gpu.t[k].copy_(torch.from_numpy(np.ascontiguousarray(cpu[k], np.float32)))
    if "id" in cpu:
        gpu.t["id"].copy_(torch.from_numpy(np.ascontiguousarray(cpu["id"], np.int32)))
    torch.cuda.synchronize()


def copyToCPU(gpu):
    """Device ParticleData -> host dict of numpy arrays (ParticleData.h:108-123)."""
    return gpu.numpy()


def to_gpu(cpu):
    p = allocateParticlesGPU(len(cpu["x"]))
    copyToGPU(cpu, p)
    return p


# ---------------------------------------------------------------------------
# OctreeData and the launchers (BHTree.cuh:38-50)
# ---------------------------------------------------------------------------
class OctreeData:
    """Opaque nbx_tree handle (tree scratch, stream) for one or more disks."""

    def __init__(self, n_bodies=None, seg_counts=None):
        h = C.c_void_p()
        if seg_counts is not None:
            arr = (C.c_int32 * len(seg_counts))(*seg_counts)
            _check(lib().nbx_tree_create_batch(C.byref(h), len(seg_counts), arr),
                   "allocateOctree(batch)")
            self.seg_counts = list(seg_counts)
        else:
            _check(lib().nbx_tree_create(C.byref(h), int(n_bodies)), "allocateOctree")
            self.seg_counts = [int(n_bodies)]
        self.h = h
        self.n_bodies = sum(self.seg_counts)

    def free(self):
        if getattr(self, "h", None) and self.h.value:
            lib().nbx_tree_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def sort_idx(self):
        torch = _torch()
        out = torch.empty(self.n_bodies, dtype=torch.int32, device="cuda")
        _check(lib().nbx_sort_idx(self.h, out.data_ptr()), "sort_idx")
        return out

    def sorted_keys(self):
        torch = _torch()
        out = torch.empty(self.n_bodies, dtype=torch.int64, device="cuda")
        _check(lib().nbx_sorted_keys(self.h, out.data_ptr()), "sorted_keys")
        return out

    def stats(self):
        ns = len(self.seg_counts)
        a = (C.c_int32 * ns)()
        b = (C.c_int32 * ns)()
        _check(lib().nbx_tree_stats(self.h, a, b), "tree_stats")
        return list(a), list(b)


def allocateOctree(n_bodies):
    return OctreeData(n_bodies)


def allocateOctreeBatch(seg_counts):
    return OctreeData(seg_counts=seg_counts)


def freeOctree(t):
    t.free()


def resetOctree(t):
    _check(lib().nbx_tree_reset(t.h), "resetOctree")


def launchBBoxKernel(t, p):
    _check(lib().nbx_bbox(t.h, C.byref(p.c())), "launchBBoxKernel")


def launchBuildKernel(t, p):
    _check(lib().nbx_build(t.h, C.byref(p.c())), "launchBuildKernel")


def launchSummarizeKernel(t):
    _check(lib().nbx_summarize(t.h), "launchSummarizeKernel")


def launchSortKernel(t):
    _check(lib().nbx_sort(t.h), "launchSortKernel")


def launchForceKernel(t, p, theta, softening, star_mass, self_acc=None):
    """a = star + tree self-gravity; self_acc (3 tensors) receives the tree part."""
    if self_acc is None:
        _check(lib().nbx_force(t.h, C.byref(p.c()), theta, softening, star_mass),
               "launchForceKernel")
    else:
        sx, sy, sz = self_acc
        _check(lib().nbx_force_ex(t.h, C.byref(p.c()), theta, softening, star_mass,
                                  sx.data_ptr(), sy.data_ptr(), sz.data_ptr()),
               "launchForceKernel")


def launchCollisionKernel(t, p, r_coll, restitution):
    _check(lib().nbx_collide(t.h, C.byref(p.c()), r_coll, restitution), "launchCollisionKernel")


def launchHalfKickKernel(p, dt_half):
    _check(lib().nbx_half_kick(C.byref(p.c()), dt_half), "launchHalfKickKernel")


def launchDriftKernel(p, dt):
    _check(lib().nbx_drift(C.byref(p.c()), dt), "launchDriftKernel")


def launchResetAccelerationKernel(p):
    _check(lib().nbx_reset_acceleration(C.byref(p.c())), "launchResetAccelerationKernel")


def launchGasDragKernel(p, drag_rate, dt, star_mass):
    _check(lib().nbx_gas_drag(C.byref(p.c()), drag_rate, dt, star_mass), "launchGasDragKernel")


def launchPerturberForceKernel(p, px, py, pz, pmass, softening):
    _check(lib().nbx_perturber_force(C.byref(p.c()), px, py, pz, pmass, softening),
           "launchPerturberForceKernel")


# ---------------------------------------------------------------------------
# Perturber (Perturber.h:27-66): host-side float arithmetic
# ---------------------------------------------------------------------------
@dataclass
class Perturber:
    enabled: bool = False
    r: float = 0.0
    mass_msun: float = 0.0
    phase0: float = 0.0
    omega: float = 0.0
    x: float = 0.0
    y: float = 0.0
    z: float = 0.0


def _f(v):
    return float(np.float32(v))


def makePerturber(enabled, mass_mjup, radius_au, phase0=0.0, star_mass=1.0):
    """makePerturber (Perturber.h:39-57) in float32 arithmetic."""
    p = Perturber(enabled=bool(enabled))
    if not p.enabled:
        return p
    f = np.float32
    p.r = _f(radius_au)
    p.mass_msun = _f(f(mass_mjup) * f(MJUP_TO_MSUN))
    p.phase0 = _f(phase0)
    r = f(p.r)
    p.omega = _f(np.sqrt(f(star_mass) / (r * r * r)))
    p.x = _f(r * np.cos(f(p.phase0)))
    p.z = _f(r * np.sin(f(p.phase0)))
    p.y = 0.0
    return p


def updatePerturber(p, sim_time):
    """updatePerturber (Perturber.h:60-66) in float32 arithmetic."""
    if not p.enabled:
        return p
    f = np.float32
    phi = f(p.phase0) + f(p.omega) * f(sim_time)
    p.x = _f(f(p.r) * np.cos(phi))
    p.z = _f(f(p.r) * np.sin(phi))
    p.y = 0.0
    return p


def particlePhysicalRadius(particle_mass_msun):
    """particlePhysicalRadius (DiskInit.cpp:10-16) in float32 arithmetic."""
    f = np.float32
    mass_kg = f(particle_mass_msun) * f(1.989e30)
    vol = mass_kg / f(3000.0)
    r_m = np.cbrt(f(3.0) * vol / (f(4.0) * f(3.14159265358979)))
    return _f(f(r_m) / f(1.496e11))


# ---------------------------------------------------------------------------
# FoF (Aggregate.h, Aggregate.cpp:26-76)
# ---------------------------------------------------------------------------
@dataclass
class Aggregate:
    id: int
    n_particles: int
    mass_msun: float
    cx: float
    cy: float
    cz: float
    vx: float
    vy: float
    vz: float
    dist_center: float
    particle_ids: list


def walk_counts(t, p, theta):
    """Mean accepted / opened records per body of the theta walk on the tree
    last summarized from p."""
    torch = _torch()
    a = torch.zeros(p.n, dtype=torch.int32, device="cuda")
    o = torch.zeros(p.n, dtype=torch.int32, device="cuda")
    _check(lib().nbx_walk_counts(t.h, theta, a.data_ptr(), o.data_ptr()), "walk_counts")
    return float(a.double().mean()), float(o.double().mean())


def search_stats(t, clear=True):
    """{pairs: colliding pairs counted, unstacked: per-lane searches that
    went on below a record without the per-lane stack} since the last clear."""
    out = (C.c_int32 * 2)()
    _check(lib().nbx_search_stats(t.h, out, int(clear)), "search_stats")
    return dict(pairs=out[0], unstacked=out[1])


def walk_stats(t, theta):
    """{acc, open: mean per body; records_per_wave; max_stack} of the theta walk."""
    out = (C.c_double * 4)()
    _check(lib().nbx_walk_stats(t.h, theta, out), "walk_stats")
    return dict(acc=out[0], open=out[1], records_per_wave=out[2], max_stack=out[3])


def walk_wave_stats(t, theta):
    """Per-wavefront work of the theta walk on the last summarized tree:
    int32 array (waves, 3) of {records evaluated, batches, deepest stack}."""
    nw = C.c_int32()
    _check(lib().nbx_walk_wave_stats(t.h, theta, None, 0, C.byref(nw)), "walk_wave_stats")
    out = np.zeros((nw.value, 3), np.int32)
    _check(lib().nbx_walk_wave_stats(t.h, theta, out.ctypes.data_as(C.c_void_p), nw.value,
                                     C.byref(nw)), "walk_wave_stats")
    return out


def fof_into(t, p, link_length, labels):
    """FoF labels into an existing int32 device tensor (no allocation).
    link_length: one value for every disk, or a list with one per disk of
    the tree."""
    if np.ndim(link_length) == 0:
        _check(lib().nbx_fof(t.h, C.byref(p.c()), link_length, labels.data_ptr()), "fof")
        return
    ll = [float(v) for v in link_length]
    if len(ll) != len(t.seg_counts):
        raise ValueError(f"{len(ll)} link lengths for a tree of {len(t.seg_counts)} disk(s)")
    arr = (C.c_float * len(ll))(*ll)
    _check(lib().nbx_fof_disks(t.h, C.byref(p.c()), arr, labels.data_ptr()), "fof")


NOALLOC_MAX_BODIES = 65536  # NBX_NOALLOC_MAX_BODIES of include/nbx.h


def fof_noalloc(p, link_length, labels, stream=None):
    """FoF labels into an existing int32 device tensor without a tree and
    without any device allocation (static scratch, at most
    NOALLOC_MAX_BODIES bodies); `stream` a torch stream or None (the null
    stream).  Calls that overlap on one GPU race."""
    s = None if stream is None else C.c_void_p(stream.cuda_stream)
    _check(lib().nbx_fof_noalloc(C.byref(p.c()), link_length, labels.data_ptr(), s),
           "fof_noalloc")


def fof_labels(t, p, link_length):
    """Device FoF partition: labels[i] = smallest index of i's group."""
    torch = _torch()
    out = torch.empty(p.n, dtype=torch.int32, device="cuda")
    _check(lib().nbx_fof(t.h, C.byref(p.c()), link_length, out.data_ptr()), "fof")
    return out


def aggregates_from_labels(labels, host, particle_mass_msun):
    """Groups of >= 2 bodies with the float32 sums of Aggregate.cpp:58-69
    (members in ascending index), ordered by smallest member index."""
    labels = np.asarray(labels)
    order = np.argsort(labels, kind="stable")
    lab_sorted = labels[order]
    starts = np.flatnonzero(np.r_[True, lab_sorted[1:] != lab_sorted[:-1]])
    ends = np.r_[starts[1:], len(labels)]
    f = np.float32
    out = []
    for s, e in zip(starts, ends):
        if e - s < 2:
            continue
        members = np.sort(order[s:e])
        acc = [f(0)] * 6
        for idx in members:
            for k, key in enumerate(("x", "y", "z", "vx", "vy", "vz")):
                acc[k] = f(acc[k] + f(host[key][idx]))
        inv = f(1.0) / f(len(members))
        c = [f(a * inv) for a in acc]
        out.append(Aggregate(
            id=len(out), n_particles=int(len(members)),
            mass_msun=_f(f(len(members)) * f(particle_mass_msun)),
            cx=_f(c[0]), cy=_f(c[1]), cz=_f(c[2]), vx=_f(c[3]), vy=_f(c[4]), vz=_f(c[5]),
            dist_center=_f(np.sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2])),
            particle_ids=[int(m) for m in members]))
    return out


def detectAggregates(t, p, link_length, particle_mass_msun):
    """detectAggregates (Aggregate.h:17-18) on device data: partition on the GPU,
    per-group float32 sums on the host as the reference does."""
    labels = fof_labels(t, p, link_length).cpu().numpy()
    return aggregates_from_labels(labels, p.numpy(), particle_mass_msun)


# ---------------------------------------------------------------------------
# Fused step (main.cu:85-140)
# ---------------------------------------------------------------------------
def step_params(dt=0.01, theta=0.5, softening=1e-3, star_mass=1.0, restitution=0.0,
                gas_drag_rate=0.0, collisions=True, use_graph=False):
    return nbx_step_params(dt, theta, softening, star_mass, restitution, gas_drag_rate,
                           int(collisions), int(use_graph))


def disk_params(r_coll=0.0, perturber=None):
    """One nbx_disk_params from r_coll and a Perturber (or None)."""
    if perturber is None or not perturber.enabled:
        return nbx_disk_params(r_coll, 0, 0.0, 0.0, 0.0, 0.0)
    return nbx_disk_params(r_coll, 1, perturber.mass_msun, perturber.r, perturber.phase0,
                           perturber.omega)


def _disk_array(dps, t):
    """ctypes array of the disk parameters of tree t: exactly one entry per
    disk (the library reads one nbx_disk_params per disk of the tree)."""
    if isinstance(dps, nbx_disk_params):
        dps = [dps]
    dps = list(dps)
    if len(dps) != len(t.seg_counts):
        raise ValueError(f"{len(dps)} disk parameter entries for a tree of "
                         f"{len(t.seg_counts)} disk(s)")
    return (nbx_disk_params * len(dps))(*dps)


def init_forces(t, p, sp, dps):
    arr = _disk_array(dps, t)
    _check(lib().nbx_init_forces(t.h, C.byref(p.c()), C.byref(sp), arr), "init_forces")


def step(t, p, sp, dps, step0, n_steps):
    arr = _disk_array(dps, t)
    _check(lib().nbx_step(t.h, C.byref(p.c()), C.byref(sp), arr, step0, n_steps), "step")


def step_concurrent(trees, ps, sp, dps_list, step0, n_steps):
    """n_steps steps of several trees (each a disk or a batch of disks)
    stepped together: every tree replays its captured step on its own
    stream, so the trees' kernels run side by side; each tree ends as
    nbx.step leaves it alone.  dps_list[i]: disk parameters of tree i (one
    per disk of that tree)."""
    n = len(trees)
    if not (len(ps) == n == len(dps_list)):
        raise ValueError("trees, particles and disk parameters differ in length")
    arrays = [_disk_array(d, t) for d, t in zip(dps_list, trees)]
    hs = (C.c_void_p * n)(*[t.h for t in trees])
    structs = [p.c() for p in ps]
    pp = (C.POINTER(nbx_particles) * n)(*[C.pointer(c) for c in structs])
    dd = (C.POINTER(nbx_disk_params) * n)(*[C.cast(a, C.POINTER(nbx_disk_params))
                                             for a in arrays])
    _check(lib().nbx_step_concurrent(hs, pp, n, C.byref(sp), dd, step0, n_steps),
           "step_concurrent")


def split_groups(counts, n_groups=2):
    """Contiguous groups of disks for step_concurrent whose body totals are as
    even as possible (the largest group total is minimal); returns lists of
    disk indices."""
    counts = [int(c) for c in counts]
    if not counts or n_groups < 1 or min(counts) <= 0:
        raise ValueError("split_groups needs at least one disk, positive body counts "
                         "and n_groups >= 1")
    k = min(n_groups, len(counts))
    best = None

    def rec(start, left, acc):
        nonlocal best
        if left == 1:
            cand = acc + [list(range(start, len(counts)))]
            worst = max(sum(counts[i] for i in g) for g in cand)
            if best is None or worst < best[0]:
                best = (worst, cand)
            return
        for end in range(start + 1, len(counts) - left + 2):
            rec(end, left - 1, acc + [list(range(start, end))])

    rec(0, k, [])
    return best[1]


def set_multipole(t, order):
    """1 = monopole (reference), 2 = monopole + quadrupole of accepted nodes."""
    _check(lib().nbx_tree_set_multipole(t.h, int(order)), "set_multipole")


PRECISION_FLOAT, PRECISION_COMPENSATED, PRECISION_FP64 = 0, 1, 2
_PRECISIONS = {"float": 0, "compensated": 1, "fp64": 2}


def set_precision(t, mode):
    """Precision of nbx_step: "float" (or False / 0), "compensated" (or True
    / 1: FP64-compensated KDK, forces FP32), "fp64" (2: the all-FP64 step,
    loaded from the float state by the next init_forces)."""
    m = _PRECISIONS[mode] if isinstance(mode, str) else int(mode)
    _check(lib().nbx_tree_set_precision(t.h, m), "set_precision")


def state64(t, p):
    """The state as a (9, n) float64 CUDA tensor [x, y, z, vx, vy, vz, ax,
    ay, az]: the FP64 state, hi + lo (compensated), or the float state."""
    import torch
    out = torch.empty((9, p.n), dtype=torch.float64, device="cuda")
    _check(lib().nbx_state64(t.h, C.byref(p.c()), C.c_void_p(out.data_ptr())), "state64")
    return out


def diagnostics(t, p, star_mass=1.0, softening=1e-3, with_self=False):
    """FP64 energy and angular-momentum sums of the state."""
    out = (C.c_double * 7)()
    _check(lib().nbx_diagnostics(t.h, C.byref(p.c()), star_mass, softening, int(with_self), out),
           "diagnostics")
    return dict(kinetic=out[0], star_potential=out[1], L=(out[2], out[3], out[4]),
                self_potential=out[5], mass=out[6],
                energy=out[0] + out[1] + out[5])


def set_radix_bits(t, bits):
    """Digit width of the key sort's radix passes (8, 9 or 11)."""
    _check(lib().nbx_tree_set_radix_bits(t.h, int(bits)), "set_radix_bits")


WAVES_LAUNCH, WAVES_LONGEST_FIRST, WAVES_XCD = 0, 1, 2  # NBX_WAVES_* of include/nbx.h


def set_wave_order(t, order):
    """Launch order of the walk's waves: WAVES_LAUNCH, WAVES_LONGEST_FIRST or
    WAVES_XCD (the default: contiguous ranges per accelerator die)."""
    _check(lib().nbx_tree_set_wave_order(t.h, int(order)), "set_wave_order")


def set_walk_shape(t, targets_per_wave):
    _check(lib().nbx_tree_set_walk_shape(t.h, int(targets_per_wave)), "set_walk_shape")


def walk_shape(t):
    v = C.c_int32()
    _check(lib().nbx_tree_get_walk_shape(t.h, C.byref(v)), "walk_shape")
    return v.value


def tune_walk(t, theta=0.5, softening=1e-3, star_mass=1.0):
    """Autotune the walk shape on the current tree; returns {targets: ms}."""
    ms = (C.c_float * 4)()
    _check(lib().nbx_tune_walk(t.h, theta, softening, star_mass, ms), "tune_walk")
    return {64: ms[0], 32: ms[1], 16: ms[2], 8: ms[3]}


def set_async(t, on):
    _check(lib().nbx_tree_set_async(t.h, int(on)), "set_async")


def sync(t):
    _check(lib().nbx_tree_sync(t.h), "sync")


def set_timing(t, on):
    _check(lib().nbx_tree_set_timing(t.h, int(on)), "set_timing")


def phase_ms(t, phase):
    v = C.c_float()
    _check(lib().nbx_phase_ms(t.h, phase.encode(), C.byref(v)), "phase_ms")
    return v.value


def probe_fp32(kind):
    """Achievable FP32 TFLOP/s: kind 0 FMA chains, 1 the walk's record mix."""
    v = C.c_double()
    _check(lib().nbx_probe_fp32(int(kind), C.byref(v)), "probe_fp32")
    return v.value


# ---------------------------------------------------------------------------
# Run support on the host (Config.h, DiskInit.h, OutputWriter.h)
# ---------------------------------------------------------------------------
def config_default():
    c = nbx_config()
    _check(lib().nbx_config_default(C.byref(c)), "config_default")
    return c


def loadConfig(path):
    """loadConfig (Config.cpp:5-55); raises RuntimeError with its messages."""
    c = nbx_config()
    _check(lib().nbx_config_load(os.fsencode(path), C.byref(c)), "loadConfig")
    return c


def config_particle_radius(cfg):
    return lib().nbx_particle_radius(C.byref(cfg))


def _host_particles(n):
    arrays = {k: np.zeros(n, np.float32) for k in FIELDS}
    arrays["id"] = np.zeros(n, np.int32)
    s = nbx_particles()
    for k in FIELDS:
        setattr(s, k, arrays[k].ctypes.data)
    s.id = arrays["id"].ctypes.data
    s.n = n
    return s, arrays


def initDisk(cfg, seed=None):
    """initDisk (DiskInit.cpp:18-67) on the host: dict of float32 arrays."""
    s, arrays = _host_particles(cfg.n_particles)
    _check(lib().nbx_init_disk(C.byref(cfg), cfg.disk_seed if seed is None else seed,
                               C.byref(s)), "initDisk")
    return arrays


class OutputWriter:
    """OutputWriter (OutputWriter.h:14-49) over nbx_writer_*."""

    def __init__(self, output_dir, n_particles, dt, output_every_n_steps):
        self.h = C.c_void_p()
        self.n = int(n_particles)
        _check(lib().nbx_writer_open(C.byref(self.h), os.fsencode(output_dir), self.n, dt,
                                     output_every_n_steps), "OutputWriter")

    def writeFrame(self, host, sim_time_T0, agg_sizes):
        s = nbx_particles()
        keep = {}
        for k in ("x", "y", "z"):
            keep[k] = np.ascontiguousarray(host[k], np.float32)
            setattr(s, k, keep[k].ctypes.data)
        s.n = len(keep["x"])
        a = np.ascontiguousarray(agg_sizes, np.uint16)
        if len(a) != self.n:
            raise RuntimeError("OutputWriter::writeFrame: agg_sizes_per_particle size mismatch")
        _check(lib().nbx_writer_frame(self.h, C.byref(s), sim_time_T0,
                                      a.ctypes.data_as(C.c_void_p)), "writeFrame")

    def writeStats(self, aggregates, frame_idx, sim_time_T0):
        rows = (nbx_aggregate * max(1, len(aggregates)))()
        for k, a in enumerate(aggregates):
            rows[k] = nbx_aggregate(a.id, a.n_particles, a.mass_msun, a.cx, a.cy, a.cz,
                                    a.vx, a.vy, a.vz, a.dist_center)
        _check(lib().nbx_writer_stats(self.h, frame_idx, sim_time_T0, len(aggregates), rows),
               "writeStats")

    def writePerturberStats(self, frame_idx, sim_time_T0, n_aggregates, px, py, pz, mass_msun,
                            omega):
        _check(lib().nbx_writer_perturber(self.h, frame_idx, sim_time_T0, n_aggregates, px, py,
                                          pz, mass_msun, omega), "writePerturberStats")

    def finalize(self, total_frames_written):
        _check(lib().nbx_writer_finalize(self.h, total_frames_written), "finalize")

    def close(self):
        if self.h.value:
            lib().nbx_writer_close(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
