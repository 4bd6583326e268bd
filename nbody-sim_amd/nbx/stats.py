"""Streaming reader of stats.csv for the analysis side.

stats.csv (written by nbx_writer_stats / nbx_writer_perturber, the format of
OutputWriter.cpp) holds one row per aggregate per output frame, plus one
row with agg_id = -1 for the perturber.  A long run's file is gigabytes, so
the reference's analysis scripts stream it with pandas in chunks of two
million rows, reading a fixed set of columns with fixed dtypes
(analyze_experiment.py:29-36, analyze_control.py:25-34), and fold rows into
per-frame summaries across chunk boundaries (analyze_experiment.py:40-62).
This module offers the same reading contract:

- :func:`read_stats` yields DataFrame chunks with exactly those columns and
  dtypes (``frame`` int32, ``sim_time_T0`` float32, ...);
- :func:`iter_frames` regroups the stream into whole frames, however the
  chunks cut them;
- :func:`per_frame_summary` computes the scripts' per-frame table (aggregate
  count and the three largest aggregates by particle count; the perturber
  row is kept apart, as in analyze_experiment.py).
"""
from __future__ import annotations

import numpy as np

# analyze_experiment.py:29-36 (the columns the analysis reads, with dtypes)
DTYPES = {
    "frame": np.int32, "sim_time_T0": np.float32, "n_aggregates": np.int32,
    "agg_id": np.int32, "n_particles": np.int32,
    "mass_Mearth": np.float32,
    "cx_AU": np.float32, "cy_AU": np.float32, "cz_AU": np.float32,
    "vx": np.float32, "vy": np.float32, "vz": np.float32,
    "dist_center_AU": np.float32,
}
# the full header written by OutputWriter.cpp (also checked by the writer tests)
HEADER = ("frame,sim_time_T0,sim_time_years,n_aggregates,agg_id,n_particles,mass_Msun,"
          "mass_Mearth,cx_AU,cy_AU,cz_AU,vx,vy,vz,dist_center_AU")
CHUNK = 2_000_000  # rows per chunk in the reference scripts
T0_TO_YEARS = 0.15915494  # analyze_experiment.py:51


def read_stats(path, chunksize=CHUNK, columns=None):
    """Yield DataFrame chunks of stats.csv with the analysis dtypes.

    columns: subset of DTYPES' keys (default: all of them, the scripts'
    usecols).  The file's header must be the writer's."""
    import pandas as pd
    with open(path, "r") as f:
        head = f.readline().strip()
    if head != HEADER:
        raise ValueError(f"{path}: not a stats.csv (header {head!r})")
    cols = list(DTYPES) if columns is None else list(columns)
    unknown = [c for c in cols if c not in DTYPES]
    if unknown:
        raise ValueError(f"unknown stats.csv columns {unknown}")
    dt = {c: DTYPES[c] for c in cols}
    yield from pd.read_csv(path, dtype=dt, usecols=cols, chunksize=chunksize)


def iter_frames(path, chunksize=CHUNK):
    """Yield (frame, DataFrame of that frame's rows) in file order, each frame
    whole even when a chunk boundary falls inside it.  Rows of a frame are
    contiguous in the file (the writer appends frame by frame)."""
    import pandas as pd
    pending = None
    for chunk in read_stats(path, chunksize):
        if pending is not None:
            chunk = pd.concat([pending, chunk], ignore_index=True)
        frames = chunk["frame"].to_numpy()
        if len(frames) == 0:
            continue
        # rows before the chunk's last frame are complete; that frame may
        # continue in the next chunk
        cut = int(np.argmax(frames == frames[-1]))
        done, pending = chunk.iloc[:cut], chunk.iloc[cut:]
        if len(done):
            starts = np.flatnonzero(np.r_[True, done["frame"].to_numpy()[1:] !=
                                          done["frame"].to_numpy()[:-1]])
            ends = np.r_[starts[1:], len(done)]
            for a, b in zip(starts, ends):
                yield int(done["frame"].iat[a]), done.iloc[a:b]
    if pending is not None and len(pending):
        yield int(pending["frame"].iat[0]), pending


def summarize_frame(rows):
    """Per-frame summary of analyze_experiment.py:40-62: aggregate count, the
    three largest aggregates by particle count (0 when absent) and the
    perturber row (None when the run has no perturber)."""
    aggs = rows[rows["agg_id"] != -1]
    pert = rows[rows["agg_id"] == -1]
    sizes = np.sort(aggs["n_particles"].to_numpy())[::-1]
    top = [int(sizes[k]) if len(sizes) > k else 0 for k in range(3)]
    t = float(rows["sim_time_T0"].iat[0])
    out = {"frame": int(rows["frame"].iat[0]), "t": t, "yrs": t * T0_TO_YEARS,
           "n_agg": int(len(aggs)), "max_mass": top[0], "sec_mass": top[1],
           "third_mass": top[2]}
    out["pert"] = None if pert.empty else {
        k: pert[k].iat[0].item() for k in ("cx_AU", "cy_AU", "cz_AU", "mass_Mearth")}
    return out


def per_frame_summary(path, chunksize=CHUNK):
    """The per-frame table of the analysis scripts as a DataFrame with
    columns frame, t, yrs, n_agg, max_mass, sec_mass, third_mass (one row
    per frame that has aggregate or perturber rows)."""
    import pandas as pd
    rows = []
    for _, fr in iter_frames(path, chunksize):
        s = summarize_frame(fr)
        s.pop("pert")
        rows.append(s)
    return pd.DataFrame(rows, columns=["frame", "t", "yrs", "n_agg", "max_mass", "sec_mass",
                                       "third_mass"])
