"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by bench.py's cpu_baseline leg, always as the checker or the timed CPU
baseline, never as part of the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_REF_PATH = os.path.join(_HERE, "_ref", "libref.so")

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")


class OrcTree(C.Structure):
    _fields_ = [
        ("n_bodies", C.c_int), ("n_total", C.c_int), ("next_node", C.c_int),
        ("pos_x", C.POINTER(C.c_float)), ("pos_y", C.POINTER(C.c_float)),
        ("pos_z", C.POINTER(C.c_float)), ("mass", C.POINTER(C.c_float)),
        ("cell_size", C.POINTER(C.c_float)), ("child", C.POINTER(C.c_int)),
        ("count", C.POINTER(C.c_int)), ("bmin", C.c_float * 3), ("bmax", C.c_float * 3),
        ("lost", C.c_int), ("max_depth", C.c_int),
    ]


class StepParams(C.Structure):
    _fields_ = [
        ("dt", C.c_float), ("theta", C.c_float), ("softening", C.c_float),
        ("star_mass", C.c_float), ("r_coll", C.c_float), ("restitution", C.c_float),
        ("gas_drag_rate", C.c_float), ("perturber_enabled", C.c_int),
        ("pert_mass_mjup", C.c_float), ("pert_radius_au", C.c_float),
        ("pert_phase0", C.c_float), ("collisions", C.c_int), ("exact_collide", C.c_int),
    ]


class Perturber(C.Structure):
    _fields_ = [("enabled", C.c_int), ("r", C.c_float), ("mass_msun", C.c_float),
                ("phase0", C.c_float), ("omega", C.c_float), ("x", C.c_float),
                ("y", C.c_float), ("z", C.c_float)]


def build():
    """Compile liboracle.so (and _ref/libref.so when the reference is present)."""
    subprocess.run(["make", "-B", "-s", "-C", _HERE, "liboracle.so"], check=True)
    if os.path.isdir("/root/reference/src"):
        subprocess.run(["make", "-B", "-s", "-C", _HERE, "ref"], check=True)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_tree_new.restype = C.POINTER(OrcTree)
        L.orc_tree_new.argtypes = [C.c_int]
        L.orc_tree_free.argtypes = [C.POINTER(OrcTree)]
        L.orc_ref_bbox.argtypes = [C.POINTER(OrcTree), _f32p, _f32p, _f32p, C.c_int]
        L.orc_ref_build.argtypes = [C.POINTER(OrcTree), _f32p, _f32p, _f32p, _f32p, C.c_int]
        L.orc_ref_summarize.argtypes = [C.POINTER(OrcTree)]
        L.orc_ref_sort.argtypes = [C.POINTER(OrcTree), _i32p]
        L.orc_ref_force.argtypes = [C.POINTER(OrcTree), _f32p, _f32p, _f32p, C.c_int,
                                    C.c_float, C.c_float, C.c_float,
                                    C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.c_void_p, C.c_int]
        L.orc_ref_collide.restype = C.c_long
        L.orc_ref_collide.argtypes = [C.POINTER(OrcTree), _f32p, _f32p, _f32p,
                                      _f32p, _f32p, _f32p, C.c_int, C.c_float, C.c_float,
                                      C.c_int, C.c_int]
        L.orc_half_kick.argtypes = [_f32p] * 6 + [C.c_float, C.c_int]
        L.orc_drift.argtypes = [_f32p] * 6 + [C.c_float, C.c_int]
        L.orc_gas_drag.argtypes = [_f32p] * 6 + [C.c_float, C.c_float, C.c_float, C.c_int]
        L.orc_perturber_force.argtypes = [_f32p] * 6 + [C.c_float] * 5 + [C.c_int]
        L.orc_make_perturber.argtypes = [C.POINTER(Perturber), C.c_int] + [C.c_float] * 4
        L.orc_update_perturber.argtypes = [C.POINTER(Perturber), C.c_float]
        L.orc_particle_physical_radius.restype = C.c_float
        L.orc_particle_physical_radius.argtypes = [C.c_float]
        L.orc_fof_n2.argtypes = [_f32p, _f32p, _f32p, C.c_int, C.c_float, _i32p]
        L.orc_fof_grid.argtypes = [_f32p, _f32p, _f32p, C.c_int, C.c_float, _i32p]
        L.orc_direct_f64.argtypes = [_f32p, _f32p, _f32p, _f32p, C.c_int, C.c_void_p, C.c_int,
                                     C.c_double, _f64p, _f64p, _f64p, C.c_int]
        L.orc_mt_root.argtypes = [_f32p, _f32p, _f32p, C.c_int, _f32p]
        L.orc_mt_keys.argtypes = [_f32p, _f32p, _f32p, C.c_int, _f32p, _u64p]
        L.orc_mt_sort.argtypes = [_u64p, C.c_int, _i32p]
        L.orc_mt_force.argtypes = [_f32p, _f32p, _f32p, _f32p, C.c_int, C.c_float, C.c_float,
                                   C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.c_int]
        L.orc_mt_force_at.argtypes = [_f32p, _f32p, _f32p, _f32p, C.c_int, C.c_float,
                                      C.c_float, C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                                      C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
        L.orc_mt_lists.restype = C.c_long
        L.orc_mt_lists.argtypes = [_f32p, _f32p, _f32p, _f32p, C.c_int, C.c_float, _i32p,
                                   C.c_void_p, C.c_void_p, C.c_void_p, C.c_long]
        L.orc_ref_init_forces.argtypes = [_f32p] * 10 + [C.c_int, C.POINTER(StepParams), C.c_int]
        L.orc_ref_steps.argtypes = [_f32p] * 10 + [C.c_int, C.POINTER(StepParams), C.c_int,
                                                  C.c_int, C.c_int]
        _lib = L
    return _lib


def ref_lib():
    """The reference's own FoF / disk-init / output-writer code, or None when not built."""
    global _ref
    if _ref is None and os.path.exists(_REF_PATH):
        R = C.CDLL(_REF_PATH)
        R.ref_detect_aggregates.restype = C.c_int
        R.ref_detect_aggregates.argtypes = [C.c_int] + [_f32p] * 6 + [
            C.c_float, C.c_float, _i32p, C.c_int, _i32p, _f32p]
        R.ref_particle_physical_radius.restype = C.c_float
        R.ref_particle_physical_radius.argtypes = [C.c_float]
        if hasattr(R, "ref_write_run"):
            R.ref_write_run.restype = C.c_int
            R.ref_write_run.argtypes = [C.c_char_p, C.c_int, C.c_float, C.c_int, C.c_int,
                                        C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                        C.c_void_p, C.c_void_p, C.c_int] + [C.c_float] * 5
        _ref = R
    return _ref


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


# ---------------------------------------------------------------------------
# Reference canonical octree
# ---------------------------------------------------------------------------
class RefTree:
    """Canonical octree of the reference (BHTree.cu), built on the CPU."""

    def __init__(self, x, y, z, m):
        self.x, self.y, self.z, self.m = _f32(x), _f32(y), _f32(z), _f32(m)
        self.n = len(self.x)
        self.t = lib().orc_tree_new(self.n)
        lib().orc_ref_bbox(self.t, self.x, self.y, self.z, self.n)
        self.lost = lib().orc_ref_build(self.t, self.x, self.y, self.z, self.m, self.n)
        lib().orc_ref_summarize(self.t)

    def __del__(self):
        if getattr(self, "t", None):
            lib().orc_tree_free(self.t)
            self.t = None

    @property
    def max_depth(self):
        return self.t.contents.max_depth

    def sort_idx(self):
        out = np.empty(self.n, np.int32)
        r = lib().orc_ref_sort(self.t, out)
        assert r == self.n, r
        return out

    def force(self, theta, softening, star_mass, counts=False, nthreads=0):
        n = self.n
        ax, ay, az = (np.zeros(n, np.float32) for _ in range(3))
        sx, sy, sz = (np.zeros(n, np.float32) for _ in range(3))
        na = np.zeros(n, np.int32) if counts else None
        no = np.zeros(n, np.int32) if counts else None
        ov = lib().orc_ref_force(self.t, self.x, self.y, self.z, n, theta, softening, star_mass,
                                 _ptr(ax), _ptr(ay), _ptr(az), _ptr(sx), _ptr(sy), _ptr(sz),
                                 _ptr(na), _ptr(no), nthreads)
        out = dict(a=np.stack([ax, ay, az], 1), self_a=np.stack([sx, sy, sz], 1), overflow=ov)
        if counts:
            out["n_acc"], out["n_open"] = na, no
        return out

    def collide(self, vx, vy, vz, r_coll, restitution, exact=True, nthreads=0):
        vx, vy, vz = _f32(vx).copy(), _f32(vy).copy(), _f32(vz).copy()
        pairs = lib().orc_ref_collide(self.t, self.x, self.y, self.z, vx, vy, vz, self.n,
                                      r_coll, restitution, int(exact), nthreads)
        return vx, vy, vz, pairs


# ---------------------------------------------------------------------------
# Integrator / perturber / FoF / direct sum
# ---------------------------------------------------------------------------
def half_kick(v, a, dt_half):
    v = [_f32(c).copy() for c in v]
    lib().orc_half_kick(v[0], v[1], v[2], _f32(a[0]), _f32(a[1]), _f32(a[2]), dt_half, len(v[0]))
    return v


def drift(x, v, dt):
    x = [_f32(c).copy() for c in x]
    lib().orc_drift(x[0], x[1], x[2], _f32(v[0]), _f32(v[1]), _f32(v[2]), dt, len(x[0]))
    return x


def gas_drag(x, v, rate, dt, star_mass):
    v = [_f32(c).copy() for c in v]
    lib().orc_gas_drag(_f32(x[0]), _f32(x[1]), _f32(x[2]), v[0], v[1], v[2], rate, dt,
                       star_mass, len(v[0]))
    return v


def perturber_force(a, x, p, pmass, softening):
    a = [_f32(c).copy() for c in a]
    lib().orc_perturber_force(a[0], a[1], a[2], _f32(x[0]), _f32(x[1]), _f32(x[2]),
                              p[0], p[1], p[2], pmass, softening, len(a[0]))
    return a


def make_perturber(enabled, mass_mjup, radius_au, phase0, star_mass):
    p = Perturber()
    lib().orc_make_perturber(C.byref(p), int(enabled), mass_mjup, radius_au, phase0, star_mass)
    return p


def update_perturber(p, t):
    lib().orc_update_perturber(C.byref(p), t)
    return p


def particle_physical_radius(m):
    return lib().orc_particle_physical_radius(m)


def fof(x, y, z, ll, method="auto"):
    x, y, z = _f32(x), _f32(y), _f32(z)
    lab = np.empty(len(x), np.int32)
    if method == "n2" or (method == "auto" and len(x) <= 4096):
        g = lib().orc_fof_n2(x, y, z, len(x), ll, lab)
    else:
        g = lib().orc_fof_grid(x, y, z, len(x), ll, lab)
    return lab, g


def direct_f64(x, y, z, m, softening, targets=None, nthreads=0):
    x, y, z, m = _f32(x), _f32(y), _f32(z), _f32(m)
    nt = len(x) if targets is None else len(targets)
    t = None if targets is None else np.ascontiguousarray(targets, np.int32)
    ax, ay, az = (np.zeros(nt, np.float64) for _ in range(3))
    lib().orc_direct_f64(x, y, z, m, len(x), _ptr(t), nt, softening, ax, ay, az, nthreads)
    return np.stack([ax, ay, az], 1)


def mt_root(x, y, z):
    out = np.zeros(4, np.float32)
    lib().orc_mt_root(_f32(x), _f32(y), _f32(z), len(x), out)
    return out


def mt_keys(x, y, z, root=None):
    x, y, z = _f32(x), _f32(y), _f32(z)
    if root is None:
        root = mt_root(x, y, z)
    keys = np.zeros(len(x), np.uint64)
    lib().orc_mt_keys(x, y, z, len(x), _f32(root), keys)
    return keys


def mt_sort(keys):
    keys = np.ascontiguousarray(keys, np.uint64)
    perm = np.zeros(len(keys), np.int32)
    lib().orc_mt_sort(keys, len(keys), perm)
    return perm


def mt_force(x, y, z, m, theta, softening, counts=False, quad=False, nthreads=0, targets=None):
    """Self-gravity of the compressed-octree walk; with `targets` (original
    indices) only those bodies are walked and rows follow the target list."""
    x, y, z, m = _f32(x), _f32(y), _f32(z), _f32(m)
    n = len(x)
    tg = None if targets is None else np.ascontiguousarray(targets, np.int32)
    nt = n if tg is None else len(tg)
    sx, sy, sz = (np.zeros(nt, np.float32) for _ in range(3))
    na = np.zeros(nt, np.int32) if counts else None
    no = np.zeros(nt, np.int32) if counts else None
    ov = lib().orc_mt_force_at(x, y, z, m, n, theta, softening, int(quad), _ptr(tg), nt,
                               _ptr(sx), _ptr(sy), _ptr(sz), _ptr(na), _ptr(no), nthreads)
    out = dict(self_a=np.stack([sx, sy, sz], 1), overflow=ov)
    if counts:
        out["n_acc"], out["n_open"] = na, no
    return out


def mt_lists(x, y, z, m, theta):
    """Interaction lists of mt_force's monopole walk: perm (sorted -> original),
    CSR offsets per target (original order) and the accepted nodes' sorted
    ranges lo, hi (inclusive)."""
    x, y, z, m = _f32(x), _f32(y), _f32(z), _f32(m)
    n = len(x)
    perm = np.zeros(n, np.int32)
    off = np.zeros(n + 1, np.int64)
    total = lib().orc_mt_lists(x, y, z, m, n, theta, perm, _ptr(off), None, None, 0)
    lo = np.zeros(max(total, 1), np.int32)
    hi = np.zeros(max(total, 1), np.int32)
    got = lib().orc_mt_lists(x, y, z, m, n, theta, perm, _ptr(off), _ptr(lo), _ptr(hi), total)
    assert got == total
    return perm, off, lo[:total], hi[:total]


def step_params(**kw):
    p = StepParams()
    defaults = dict(dt=0.01, theta=0.5, softening=1e-3, star_mass=1.0, r_coll=0.0,
                    restitution=0.0, gas_drag_rate=0.0, perturber_enabled=0,
                    pert_mass_mjup=0.0, pert_radius_au=5.2, pert_phase0=0.0, collisions=1,
                    exact_collide=1)
    defaults.update(kw)
    for k, v in defaults.items():
        setattr(p, k, v)
    return p


def ref_steps(state, m, params, step0, steps, init_forces=False, nthreads=0):
    """Run the reference step loop (main.cu:85-140) on copies of `state`
    (dict of x,y,z,vx,vy,vz,ax,ay,az float32 arrays).  Returns the new state."""
    s = {k: _f32(state[k]).copy() for k in ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az")}
    m = _f32(m)
    n = len(m)
    args = [s[k] for k in ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az")] + [m, n,
                                                                                C.byref(params)]
    if init_forces:
        lib().orc_ref_init_forces(*args, nthreads)
    if steps:
        lib().orc_ref_steps(*args, step0, steps, nthreads)
    return s
