"""Seeded synthetic workloads for the five benchmark configurations.

Velocity fields are sums of random Fourier modes: |k| log-uniform, random direction, uniform
phase, amplitude a ~ |k|^p, amplitude vector perpendicular to k (divergence-free).  The nearest
precedent in the reference is docs/generate_assets/generate_assets.jl:24-46.  All draws come
from numpy's PCG64 generator with the seeds listed per configuration.
"""
from __future__ import annotations

import numpy as np

BASE_SEED = 20261202


def _flat_field(rng, pts, n_modes, kmin, kmax, power, chunk=32):
    """u(x) = sum_m a_m e_m cos(k_m . x + phi_m), e_m unit and perpendicular to k_m."""
    D, N = pts.shape
    kmag = np.exp(rng.uniform(np.log(kmin), np.log(kmax), n_modes))
    if D == 2:
        th = rng.uniform(0, 2 * np.pi, n_modes)
        khat = np.stack([np.cos(th), np.sin(th)])
        ehat = np.stack([-np.sin(th), np.cos(th)])
    else:
        v = rng.normal(size=(3, n_modes))
        khat = v / np.linalg.norm(v, axis=0)
        r = rng.normal(size=(3, n_modes))
        r -= khat * np.sum(r * khat, axis=0)
        ehat = r / np.linalg.norm(r, axis=0)
    amp = kmag ** power
    phase = rng.uniform(0, 2 * np.pi, n_modes)
    k = khat * kmag
    u = np.zeros((D, N))
    for m0 in range(0, n_modes, chunk):
        m1 = min(n_modes, m0 + chunk)
        arg = k[:, m0:m1].T @ pts + phase[m0:m1, None]
        c = np.cos(arg) * amp[m0:m1, None]
        u += ehat[:, m0:m1] @ c
    return u


def config1(seed=BASE_SEED + 1, N=4096):
    """2D, N=4096 uniform in [0,1]^2, 256 modes, |k| in [2pi, 2pi*256], a ~ |k|^(-5/6);
    L2 over 33 log edges 1e-3..1.414."""
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, 1, (2, N))
    u = _flat_field(rng, x, 256, 2 * np.pi, 2 * np.pi * 256, -5 / 6)
    return x, u, np.geomspace(1e-3, 1.414, 33)


def config2(seed=BASE_SEED + 2, N=65536):
    """2D, N=65536, same field recipe; 65 log edges 1e-4..1.414."""
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, 1, (2, N))
    u = _flat_field(rng, x, 256, 2 * np.pi, 2 * np.pi * 256, -5 / 6)
    return x, u, np.geomspace(1e-4, 1.414, 65)


def config3(seed=BASE_SEED + 3, N=131072):
    """3D, N=131072 uniform in [0,1]^3, 512 solenoidal modes, |k| in [2pi, 2pi*128], a ~ |k|^(-11/6);
    49 log distance edges 1e-4..1.732 x 129 linear value edges over [-4s, 4s], s = sqrt(2) rms(u_x)."""
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, 1, (3, N))
    u = _flat_field(rng, x, 512, 2 * np.pi, 2 * np.pi * 128, -11 / 6)
    s = np.sqrt(2.0) * np.sqrt(np.mean(u[0] ** 2))
    return x, u, np.geomspace(1e-4, 1.732, 49), np.linspace(-4 * s, 4 * s, 129)


def sphere_points(rng, N, n_tracks=512, per_track=256, spacing=1e-3, jitter=1e-4):
    """Half uniform on the unit sphere, half on random great-circle tracks; (lon, lat) degrees."""
    n_uni = N - n_tracks * per_track if N >= 2 * n_tracks * per_track else N // 2
    n_trk = N - n_uni
    v = rng.normal(size=(3, n_uni))
    v /= np.linalg.norm(v, axis=0)
    pts = [v]
    tracks = max(1, (n_trk + per_track - 1) // per_track)
    left = n_trk
    for _ in range(tracks):
        m = min(per_track, left)
        if m <= 0:
            break
        a = rng.normal(size=3)
        a /= np.linalg.norm(a)
        b = rng.normal(size=3)
        b -= a * (a @ b)
        b /= np.linalg.norm(b)
        t0 = rng.uniform(0, 2 * np.pi)
        t = t0 + spacing * np.arange(m)
        p = np.outer(a, np.cos(t)) + np.outer(b, np.sin(t))
        p += jitter * rng.normal(size=p.shape)
        p /= np.linalg.norm(p, axis=0)
        pts.append(p)
        left -= m
    P = np.concatenate(pts, axis=1)
    lon = np.degrees(np.arctan2(P[1], P[0]))
    lat = np.degrees(np.arcsin(np.clip(P[2], -1, 1)))
    return np.stack([lon, lat]), P


def config4(seed=BASE_SEED + 4, N=262144):
    """Unit sphere, N=262144 (half uniform, half on 512 tracks x 256 points at 1e-3 rad spacing with
    1e-4 rad jitter), 128 R^3 modes |k| in [1, 256], a ~ |k|^(-5/6), projected on (east, north);
    65 log distance edges 1e-5..pi, 97 linear value edges over [-4s, 4s] (s as config 3)."""
    rng = np.random.default_rng(seed)
    x, P = sphere_points(rng, N)
    U = _flat_field(rng, P, 128, 1.0, 256.0, -5 / 6)
    lon, lat = np.radians(x[0]), np.radians(x[1])
    E = np.stack([-np.sin(lon), np.cos(lon), np.zeros_like(lon)])
    Nn = np.stack([-np.sin(lat) * np.cos(lon), -np.sin(lat) * np.sin(lon), np.cos(lat)])
    u = np.stack([np.sum(U * E, axis=0), np.sum(U * Nn, axis=0)])
    s = np.sqrt(2.0) * np.sqrt(np.mean(u[0] ** 2))
    return x, u, np.geomspace(1e-5, np.pi, 65), np.linspace(-4 * s, 4 * s, 97)


def config5(seed=20261207, n_slices=256, nmin=2000, nmax=40000):
    """256 2D slices, N_s = round(exp(U[ln 2000, ln 40000])), each with its own 256-mode field
    (seed seed*1000 + s); 49 log distance edges 1e-4..1.414, 65 value edges over [-4s_s, 4s_s]."""
    rng = np.random.default_rng(seed)
    Ns = np.round(np.exp(rng.uniform(np.log(nmin), np.log(nmax), n_slices))).astype(np.int64)
    xs, us, vbins = [], [], []
    for sidx, n in enumerate(Ns):
        r = np.random.default_rng(seed * 1000 + sidx)
        x = r.uniform(0, 1, (2, int(n)))
        u = _flat_field(r, x, 256, 2 * np.pi, 2 * np.pi * 256, -5 / 6)
        s = np.sqrt(2.0) * np.sqrt(np.mean(u[0] ** 2))
        xs.append(x)
        us.append(u)
        vbins.append(np.linspace(-4 * s, 4 * s, 65))
    return xs, us, np.geomspace(1e-4, 1.414, 49), vbins


def total_pairs(Ns) -> int:
    return int(sum(int(n) * (int(n) - 1) // 2 for n in Ns))
