"""Seeded synthetic protoplanetary disk (the benchmark workloads' generator).

Follows the reference initialiser's laws (DiskInit.cpp:37-67): surface density
Sigma ~ r^-1.5 sampled by the inverse CDF r = (u (sqrt(rmax)-sqrt(rmin)) +
sqrt(rmin))^2, uniform azimuth in the XZ plane, Gaussian thickness along Y,
prograde circular Keplerian speed v = sqrt(M/r) with (-sin phi, 0, cos phi)
direction.  The 1 % isotropic velocity noise of the reference is replaced by
epicyclic eccentricity / inclination draws (Rayleigh e, i), and the random
stream is numpy's counter-based Philox so that a seed fixes the disk on any
host.
"""
from __future__ import annotations

import numpy as np

MJUP_TO_MSUN = 9.5479e-4  # Perturber.h:25


def make_disk(n, disk_mass, r_min=0.5, r_max=4.0, seed=0, h_factor=0.01, e_sigma=0.01,
              i_sigma=0.005, star_mass=1.0):
    """Return dict of float32 SoA arrays x,y,z,vx,vy,vz,ax,ay,az,mass and int32 id."""
    rng = np.random.Generator(np.random.Philox(seed))
    u = rng.random(n)
    sr = u * (np.sqrt(r_max) - np.sqrt(r_min)) + np.sqrt(r_min)
    r = sr * sr
    phi = rng.random(n) * 2.0 * np.pi
    y = rng.standard_normal(n) * (h_factor * r)
    vc = np.sqrt(star_mass / r)
    e = rng.rayleigh(e_sigma, n) if e_sigma > 0 else np.zeros(n)
    psi = rng.random(n) * 2.0 * np.pi
    inc = rng.rayleigh(i_sigma, n) if i_sigma > 0 else np.zeros(n)
    om = rng.random(n) * 2.0 * np.pi
    v_r = e * vc * np.sin(psi)
    v_t = vc * (1.0 + 0.5 * e * np.cos(psi))
    v_y = inc * vc * np.cos(om)
    c, s = np.cos(phi), np.sin(phi)
    f = np.float32
    out = dict(
        x=(r * c).astype(f), y=y.astype(f), z=(r * s).astype(f),
        vx=(v_r * c - v_t * s).astype(f), vy=v_y.astype(f), vz=(v_r * s + v_t * c).astype(f),
        ax=np.zeros(n, f), ay=np.zeros(n, f), az=np.zeros(n, f),
        mass=np.full(n, np.float32(disk_mass / n), f),
        id=np.arange(n, dtype=np.int32),
    )
    return out
