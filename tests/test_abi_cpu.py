"""CPU tests of the C ABI library and the host-side logic (no GPU calls).

libnbx.so must load and export every entry point include/nbx.h declares;
argument errors must be reported before any device work; the host-side
float32 restatements (perturber orbit, particle radius) must agree with the
oracle's C versions.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

import nbx
import pyoracle as O

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(REPO, "include", "nbx.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nbx_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(nbx.LIB_PATH)
    declared = _declared()
    assert len(declared) >= 30
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, missing
    assert sorted(nbx.EXPORTS) == declared


def test_null_arguments_fail_before_device_work():
    L = nbx.lib()
    assert L.nbx_force(None, None, 0.5, 1e-3, 1.0) == -1
    assert b"null tree" in L.nbx_last_error()
    assert L.nbx_half_kick(None, 0.005) == -1
    assert b"invalid particle" in L.nbx_last_error()
    assert L.nbx_tree_create(None, 10) == -1
    h = C.c_void_p()
    assert L.nbx_tree_create_batch(C.byref(h), 0, None) == -1
    counts = (C.c_int32 * 2)(10, 0)
    assert L.nbx_tree_create_batch(C.byref(h), 2, counts) == -1
    assert b"disk 1 has 0 bodies" in L.nbx_last_error()
    too_many = (C.c_int32 * 17)(*([1] * 17))
    assert L.nbx_tree_create_batch(C.byref(h), 17, too_many) == -1
    assert L.nbx_tree_destroy(None) == 0
    assert L.nbx_phase_ms(None, b"walk", None) == -1
    assert L.nbx_step_concurrent(None, None, 1, None, None, 0, 1) == -1
    assert b"null arguments" in L.nbx_last_error()


def test_split_groups_contiguous_and_even():
    """Contiguous disk groups with the smallest largest body total."""
    sweep = [16384, 32768, 65536, 98304, 131072, 262144]
    assert nbx.split_groups(sweep, 2) == [[0, 1, 2, 3, 4], [5]]
    assert nbx.split_groups(sweep, 1) == [list(range(6))]
    assert nbx.split_groups([5, 5, 5, 5], 2) == [[0, 1], [2, 3]]
    assert nbx.split_groups([7], 3) == [[0]]
    assert nbx.split_groups([1, 2, 3], 5) == [[0], [1], [2]]
    for bad in (([], 2), ([4, 0], 2), ([3], 0)):
        with pytest.raises(ValueError):
            nbx.split_groups(*bad)
    # brute force over every contiguous split of a ragged list
    counts = [9, 1, 4, 7, 2, 8, 3]
    for k in (2, 3):
        g = nbx.split_groups(counts, k)
        assert sum(g, []) == list(range(len(counts))) and len(g) == k
        import itertools
        best = min(max(sum(counts[a:b]) for a, b in zip((0,) + c, c + (len(counts),)))
                   for c in itertools.combinations(range(1, len(counts)), k - 1))
        assert max(sum(counts[i] for i in x) for x in g) == best


def test_missing_library_fails_loudly(monkeypatch):
    monkeypatch.setattr(nbx, "_lib", None)
    monkeypatch.setattr(nbx, "LIB_PATH", "/nonexistent/libnbx.so")
    with pytest.raises(RuntimeError, match="not built"):
        nbx.lib()


@pytest.mark.parametrize("mass_mjup,radius,phase0", [(1.0, 10.0, 0.0), (10.0, 5.2, 0.3),
                                                     (0.3, 20.0, 1.0)])
def test_host_perturber_matches_oracle(mass_mjup, radius, phase0):
    p = nbx.makePerturber(True, mass_mjup, radius, phase0, 1.0)
    q = O.make_perturber(1, mass_mjup, radius, phase0, 1.0)
    assert np.float32(p.mass_msun) == np.float32(q.mass_msun)
    assert np.float32(p.omega) == np.float32(q.omega)
    for t in (0.0, 0.01, 1.37, 250.0):
        nbx.updatePerturber(p, t)
        O.update_perturber(q, t)
        # numpy's float32 cos/sin and glibc cosf/sinf may differ by one ulp
        assert abs(p.x - q.x) <= 2 * np.spacing(np.float32(radius))
        assert abs(p.z - q.z) <= 2 * np.spacing(np.float32(radius))


def test_host_particle_radius_matches_oracle():
    for m in (3.003e-7, 1e-5 / 4096, 3e-5 / 131072):
        a = nbx.particlePhysicalRadius(m)
        b = O.particle_physical_radius(m)
        assert abs(a - b) <= 2 * np.spacing(np.float32(b))


def test_aggregates_from_labels_groups_and_order():
    labels = np.array([0, 0, 2, 3, 2, 5, 0], np.int32)
    host = {k: np.arange(7, dtype=np.float32) for k in ("x", "y", "z", "vx", "vy", "vz")}
    aggs = nbx.aggregates_from_labels(labels, host, 1e-6)
    assert [a.n_particles for a in aggs] == [3, 2]
    assert aggs[0].particle_ids == [0, 1, 6] and aggs[1].particle_ids == [2, 4]
    assert abs(aggs[0].cx - 7 / 3) < 1e-6 and abs(aggs[1].cx - 3.0) < 1e-6


def test_step_params_layout_matches_header():
    """The ctypes structs mirror include/nbx.h field for field."""
    text = open(os.path.join(REPO, "include", "nbx.h")).read()
    for struct, cls in (("nbx_step_params", nbx.nbx_step_params),
                        ("nbx_disk_params", nbx.nbx_disk_params),
                        ("nbx_particles", nbx.nbx_particles)):
        body = re.search(r"typedef struct %s \{(.*?)\} %s;" % (struct, struct), text, re.S)
        assert body, struct
        fields = re.findall(r"\b([a-z_0-9]+)\s*[,;]", re.sub(r"/\*.*?\*/", "", body.group(1)))
        assert fields == [f[0] for f in cls._fields_], (struct, fields)


def test_disk_params_length_checked_before_device_work():
    """The library reads one nbx_disk_params per disk of a tree: the Python
    wrappers refuse a list of another length before any call (a shorter
    array would be read past its end)."""
    class FakeTree:
        seg_counts = [100, 200, 300]
        h = None
    dp = nbx.disk_params(1e-3)
    for fn in (lambda t, d: nbx.init_forces(t, None, None, d),
               lambda t, d: nbx.step(t, None, None, d, 0, 1)):
        with pytest.raises(ValueError, match="3 disk"):
            fn(FakeTree(), dp)
        with pytest.raises(ValueError, match="2 disk parameter entries"):
            fn(FakeTree(), [dp, dp])
    with pytest.raises(ValueError, match="1 disk parameter entries"):
        nbx.step_concurrent([FakeTree()], [None], None, [[dp]], 0, 1)
    one = FakeTree()
    one.seg_counts = [10]
    assert len(nbx._disk_array(dp, one)) == 1
    assert len(nbx._disk_array([dp, dp, dp], FakeTree())) == 3


def test_fof_noalloc_rejects_oversized_input_before_device_work():
    """The static-scratch FoF takes at most NBX_NOALLOC_MAX_BODIES bodies:
    a larger count is an argument error, reported without touching a
    device (the pointers are never dereferenced)."""
    L = nbx.lib()
    text = open(os.path.join(REPO, "include", "nbx.h")).read()
    assert f"#define NBX_NOALLOC_MAX_BODIES {nbx.NOALLOC_MAX_BODIES}" in text
    s = nbx.nbx_particles()
    for k in nbx.FIELDS:
        setattr(s, k, 16)  # non-null stand-ins, never read
    s.id = 16
    s.n = nbx.NOALLOC_MAX_BODIES + 1
    assert L.nbx_fof_noalloc(C.byref(s), 1e-3, C.c_void_p(16), None) == -1
    assert b"exceed" in L.nbx_last_error()
    assert L.nbx_fof_noalloc(None, 1e-3, C.c_void_p(16), None) == -1
    s.n = 100
    assert L.nbx_fof_noalloc(C.byref(s), 1e-3, None, None) == -1
    assert b"null labels" in L.nbx_last_error()
