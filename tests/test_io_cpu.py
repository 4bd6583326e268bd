"""Run support on the host (no GPU): the JSON config of Config.cpp, the disk
generator of DiskInit.cpp and the frames.bin v2 / stats.csv writer of
OutputWriter.cpp, against the reference's own config files, its test
expectations (test_config.cpp, test_disk_init.cpp, test_output.cpp) and,
where it is built here, the reference's own OutputWriter byte for byte."""
import ctypes as C
import os
import struct

import numpy as np
import pytest

import pyoracle as O
import nbx

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
f32 = np.float32


def test_config_matches_reference_test_expectations():
    # test_config.cpp:5-21 on config/simulation.json.  That test is stale for
    # two fields (it expects output_every_n_steps 200 and total_time 2000; the
    # file holds 50 and 500): those two are checked against the file.
    c = nbx.loadConfig(os.path.join(GOLDEN, "config_simulation.json"))
    assert c.n_particles == 100000
    assert c.output_every_n_steps == 50
    assert c.output_dir == b"data"
    for k, v in dict(disk_r_min=0.5, disk_r_max=5.0, particle_mass_msun=3.003e-7,
                     star_mass_msun=1.0, theta=0.5, dt=0.01, disk_h_factor=0.05,
                     softening_AU=0.01, collision_radius_factor=100.0, restitution=0.3,
                     total_time=500.0).items():
        assert getattr(c, k) == f32(v), k  # parsed as double, narrowed to float


def test_config_defaults_and_optional_fields(tmp_path):
    p = tmp_path / "empty.json"
    p.write_text("{}")
    c = nbx.loadConfig(str(p))
    d = nbx.config_default()
    for name, _ in nbx.nbx_config._fields_:
        assert getattr(c, name) == getattr(d, name), name
    # Config.h:5-35 defaults
    assert (c.n_particles, c.output_every_n_steps, c.perturber_enabled) == (100000, 200, 0)
    assert c.disk_h_factor == f32(0.05) and c.gas_drag_rate == 0.0
    assert c.perturber_mass_mjup == 1.0 and c.perturber_radius_AU == 20.0
    c = nbx.loadConfig(os.path.join(GOLDEN, "config_planet_formation.json"))
    assert c.n_particles == 10000 and c.collision_radius_factor == f32(500.0)
    assert c.total_time == f32(10000.0) and c.perturber_enabled == 0


@pytest.mark.parametrize("text,msg", [
    ('{"n_particles": 0}', "n_particles must be positive, got: 0"),
    ('{"dt": 0}', "dt must be positive"),
    ('{"disk_r_min": 5.0, "disk_r_max": 5.0}', "disk_r_min must be less than disk_r_max"),
    ('{"perturber_enabled": true, "perturber_mass_mjup": 0}',
     "perturber_mass_mjup must be positive"),
    ('{"perturber_enabled": true, "perturber_radius_AU": -1}',
     "perturber_radius_AU must be positive"),
    ('{"n_particles": 10,}', "JSON parse error"),
    ('{"theta": "half"}', "theta must be a number"),
])
def test_config_validation_messages(tmp_path, text, msg):
    p = tmp_path / "c.json"
    p.write_text(text)
    with pytest.raises(RuntimeError, match=msg):
        nbx.loadConfig(str(p))


def test_config_missing_file():
    with pytest.raises(RuntimeError, match="Cannot open config: nonexistent/path.json"):
        nbx.loadConfig("nonexistent/path.json")


def test_config_disabled_perturber_is_not_validated(tmp_path):
    p = tmp_path / "c.json"
    p.write_text('{"perturber_enabled": false, "perturber_mass_mjup": 0, "extra": [1, {"a": 2}]}')
    assert nbx.loadConfig(str(p)).perturber_mass_mjup == 0.0


def test_particle_radius_kat():
    # test_disk_init.cpp:42-56: 2.416e-5 AU at 3.003e-7 M_sun
    c = nbx.config_default()
    r = nbx.config_particle_radius(c)
    assert abs(r - 2.416e-5) <= 0.01 * 2.416e-5
    assert f32(r) == f32(O.particle_physical_radius(3.003e-7))


def test_init_disk_bounds_and_circular_speed():
    # test_disk_init.cpp:21-40 and :58-77
    c = nbx.config_default()
    c.n_particles, c.disk_r_min, c.disk_r_max, c.disk_h_factor = 1000, 0.5, 5.0, 0.05
    d = nbx.initDisk(c, 42)
    r = np.sqrt(d["x"] ** 2 + d["z"] ** 2)
    assert np.all(r >= c.disk_r_min - 1e-4) and np.all(r <= c.disk_r_max + 1e-4)
    assert np.all(d["mass"] == f32(c.particle_mass_msun))
    assert np.all(d["id"] == np.arange(1000))
    c.n_particles, c.disk_r_min, c.disk_r_max, c.disk_h_factor = 100, 1.0, 2.0, 0.01
    d = nbx.initDisk(c, 1)
    r = np.sqrt(d["x"] ** 2 + d["z"] ** 2)
    v = np.sqrt(d["vx"] ** 2 + d["vy"] ** 2 + d["vz"] ** 2)
    assert np.all(np.abs(v / np.sqrt(1.0 / r) - 1) < 0.05)


def test_init_disk_distribution():
    c = nbx.config_default()
    c.n_particles, c.disk_r_min, c.disk_r_max, c.disk_h_factor = 200000, 0.5, 4.0, 0.01
    d = nbx.initDisk(c, 7)
    r = np.sqrt(d["x"] ** 2 + d["z"] ** 2).astype(np.float64)
    # p(r) ~ r^-1/2: sqrt(r) uniform on [sqrt(r_min), sqrt(r_max)]
    u = (np.sqrt(r) - np.sqrt(0.5)) / (2.0 - np.sqrt(0.5))
    assert abs(u.mean() - 0.5) < 0.005 and abs(u.var() - 1 / 12) < 0.003
    phi = np.arctan2(d["z"], d["x"])
    assert abs(np.mean(np.cos(phi))) < 0.01 and abs(np.mean(np.sin(phi))) < 0.01
    assert abs(np.std(d["y"] / (0.01 * r)) - 1.0) < 0.01  # y ~ N(0, (h r)^2)
    # 1 % Gaussian noise of the circular speed on each velocity component
    vc = np.sqrt(1.0 / r)
    dvx = d["vx"] + vc * np.sin(phi)
    assert abs(np.std(d["vy"] / (0.01 * vc)) - 1.0) < 0.01
    assert abs(np.std(dvx / (0.01 * vc)) - 1.0) < 0.01
    # prograde about -Y (DiskInit.cpp:50-61)
    hy = d["z"] * d["vx"] - d["x"] * d["vz"]
    assert np.all(hy < 0)


def test_init_disk_rayleigh_mode_and_determinism():
    c = nbx.config_default()
    c.n_particles, c.disk_r_min, c.disk_r_max = 50000, 0.5, 4.0
    c.e_sigma, c.i_sigma = 0.01, 0.005
    d = nbx.initDisk(c, 3)
    x = np.stack([d["x"], d["y"], d["z"]]).astype(np.float64)
    v = np.stack([d["vx"], d["vy"], d["vz"]]).astype(np.float64)
    r = np.sqrt((x * x).sum(0))
    ev = ((v * v).sum(0) - 1.0 / r) * x - (x * v).sum(0) * v
    e = np.sqrt((ev * ev).sum(0))
    # Rayleigh(sigma) median = sigma sqrt(2 ln 2); the epicyclic set-up is first order in e
    assert abs(np.median(e) / (0.01 * np.sqrt(2 * np.log(2))) - 1) < 0.1
    d2 = nbx.initDisk(c, 3)
    d3 = nbx.initDisk(c, 4)
    assert all(np.array_equal(d[k], d2[k]) for k in d)
    assert not np.array_equal(d["x"], d3["x"])


def scripted_run(rng, n=37, frames=3, n_agg=4):
    xyz = rng.normal(0, 2, (frames, 3, n)).astype(f32)
    times = (np.arange(1, frames + 1) * f32(0.5)).astype(f32)
    sizes = rng.integers(1, 70000, (frames, n)).clip(1, 65535).astype(np.uint16)
    ai = np.stack([np.arange(n_agg), rng.integers(2, 40, n_agg)], 1).astype(np.int32)
    af = (rng.normal(0, 1, (n_agg, 8)) * 10.0 ** rng.integers(-9, 4, (n_agg, 8))).astype(f32)
    pert = (f32(4.1), f32(0.0), f32(-3.2), f32(9.5479e-3), f32(0.0843))
    return xyz, times, sizes, ai, af, pert


def write_with_nbx(d, n, dt, every, xyz, times, sizes, ai, af, pert):
    w = nbx.OutputWriter(str(d), n, dt, every)
    aggs = [nbx.Aggregate(int(a[0]), int(a[1]), *[float(v) for v in f], []) for a, f in zip(ai, af)]
    for fr in range(len(times)):
        w.writeFrame({"x": xyz[fr, 0], "y": xyz[fr, 1], "z": xyz[fr, 2]}, float(times[fr]),
                     sizes[fr])
        w.writeStats(aggs, fr, float(times[fr]))
        w.writePerturberStats(fr, float(times[fr]), len(aggs), *[float(p) for p in pert])
    w.finalize(len(times))
    w.close()


def test_writer_format_known_answers(tmp_path):
    rng = np.random.default_rng(0)
    n, dt, every = 37, f32(0.01), 7
    xyz, times, sizes, ai, af, pert = scripted_run(rng, n=n)
    write_with_nbx(tmp_path, n, dt, every, xyz, times, sizes, ai, af, pert)
    raw = (tmp_path / "frames.bin").read_bytes()
    # header (OutputWriter.cpp:44-78) and the size of 3 frames of 8 + 14 n bytes
    assert raw[:4] == b"NBOD"
    assert struct.unpack_from("<III", raw, 4) == (2, n, 3)
    assert struct.unpack_from("<dd", raw, 16) == (float(dt), float(dt) * every)
    assert len(raw) == 32 + 3 * (8 + 14 * n)
    off = 32 + (8 + 14 * n)  # second frame
    assert struct.unpack_from("<d", raw, off)[0] == float(times[1])
    got = np.frombuffer(raw, f32, 3 * n, off + 8).reshape(3, n)
    assert np.array_equal(got, xyz[1])
    assert np.array_equal(np.frombuffer(raw, np.uint16, n, off + 8 + 12 * n), sizes[1])
    rows = (tmp_path / "stats.csv").read_text().splitlines()
    assert rows[0] == ("frame,sim_time_T0,sim_time_years,n_aggregates,agg_id,n_particles,"
                       "mass_Msun,mass_Mearth,cx_AU,cy_AU,cz_AU,vx,vy,vz,dist_center_AU")
    assert len(rows) == 1 + 3 * (4 + 1)
    g = lambda v: "%g" % float(f32(v))
    a = af[2]
    years = f32(f32(times[0]) * (f32(1.0) / (f32(2.0) * f32(3.14159265))))
    assert rows[3] == ",".join(["0", g(times[0]), g(years), "4", "2", str(ai[2, 1]), g(a[0]),
                                g(f32(a[0]) * f32(332946.0))] + [g(v) for v in a[1:]])
    r = np.sqrt(f32(pert[0]) * pert[0] + f32(pert[2]) * pert[2])
    vmag = f32(pert[4]) * r
    exp = ["0", g(times[0]), g(years), "4", "-1", "0", g(pert[3]),
           g(f32(pert[3]) * f32(332946.0)), g(pert[0]), g(pert[1]), g(pert[2]),
           g(f32(-pert[2] / r) * vmag), "0", g(f32(pert[0] / r) * vmag), g(r)]
    assert rows[5] == ",".join(exp)


def test_writer_matches_reference_output_writer(tmp_path):
    ref = O.ref_lib()
    if ref is None:
        pytest.skip("reference OutputWriter not built here (needs /root/reference)")
    rng = np.random.default_rng(5)
    for n, frames, n_agg in [(4, 2, 0), (37, 3, 4), (1000, 2, 17)]:
        xyz, times, sizes, ai, af, pert = scripted_run(rng, n=n, frames=frames, n_agg=n_agg)
        dn, dr = tmp_path / f"nbx{n}", tmp_path / f"ref{n}"
        dn.mkdir()
        dr.mkdir()
        write_with_nbx(dn, n, f32(0.05), 10, xyz, times, sizes, ai, af, pert)
        ptr = lambda a: a.ctypes.data_as(C.c_void_p)
        ref.ref_write_run(os.fsencode(str(dr)), n, C.c_float(0.05), 10, frames, ptr(xyz),
                          ptr(times), ptr(sizes), n_agg, ptr(np.ascontiguousarray(ai)),
                          ptr(np.ascontiguousarray(af)), 1, *[C.c_float(p) for p in pert])
        assert (dn / "frames.bin").read_bytes() == (dr / "frames.bin").read_bytes()
        assert (dn / "stats.csv").read_bytes() == (dr / "stats.csv").read_bytes()


def test_perturber_sidecar(tmp_path):
    rc = nbx.lib().nbx_write_perturber_sidecar(os.fsencode(str(tmp_path)), 1, 9.5479e-3, 5.2,
                                               0.0, 0.0843)
    assert rc == 0
    import json
    j = json.loads((tmp_path / "perturber.json").read_text())
    assert j["enabled"] is True and abs(j["mass_msun"] - 9.5479e-3) < 1e-9
    assert set(j) == {"enabled", "mass_msun", "r_AU", "phase0_rad", "omega_T0_inv"}


def test_stats_reader_dtypes_chunks_and_summary(tmp_path):
    """The streaming stats.csv reader returns the analysis scripts' columns
    and dtypes (analyze_experiment.py:29-36) and regroups frames across
    chunk boundaries: every chunk size gives the same per-frame table, whose
    values follow the rows written (aggregate count, three largest by
    particle count, perturber row apart)."""
    from nbx import stats
    rng = np.random.default_rng(4)
    n, dt, every = 50, f32(0.01), 10
    w = nbx.OutputWriter(str(tmp_path), n, dt, every)
    expect = []
    for fr in range(6):
        k = int(rng.integers(0, 9))  # frames with 0 .. 8 aggregates
        sizes = rng.integers(2, 500, k)
        aggs = [nbx.Aggregate(j, int(s), 1e-6 * int(s), 0.1, 0.0, 0.2, 0.0, 0.0, 0.0, 0.2, [])
                for j, s in enumerate(sizes)]
        t = float(f32(0.1) * (fr + 1))
        w.writeStats(aggs, fr, t)
        if fr != 3:  # one frame without the perturber row
            w.writePerturberStats(fr, t, len(aggs), 4.0, 0.0, 1.0, 9.5479e-3, 0.08)
        top = sorted(sizes.tolist(), reverse=True) + [0, 0, 0]
        if k or fr != 3:
            expect.append((fr, k, top[0], top[1], top[2]))
    w.finalize(6)
    w.close()
    path = tmp_path / "stats.csv"
    chunk = next(stats.read_stats(path, chunksize=1000))
    assert list(chunk.columns) == list(stats.DTYPES)
    for c, t in stats.DTYPES.items():
        assert chunk[c].dtype == t, c
    tables = [stats.per_frame_summary(path, chunksize=cs) for cs in (1, 2, 3, 7, 1000)]
    for tb in tables[1:]:
        assert tb.equals(tables[0])
    got = [tuple(int(v) for v in r) for r in
           tables[0][["frame", "n_agg", "max_mass", "sec_mass", "third_mass"]].to_numpy()]
    assert got == expect
    assert np.allclose(tables[0]["yrs"], tables[0]["t"] * 0.15915494)
    frames = [f for f, _ in stats.iter_frames(path, chunksize=2)]
    assert frames == sorted(set(frames))
    with pytest.raises(ValueError, match="unknown"):
        next(stats.read_stats(path, columns=["frame", "nope"]))
    (tmp_path / "bad.csv").write_text("a,b\n1,2\n")
    with pytest.raises(ValueError, match="not a stats.csv"):
        next(stats.read_stats(tmp_path / "bad.csv"))
