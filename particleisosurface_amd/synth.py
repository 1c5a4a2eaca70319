"""Deterministic synthetic particle frames for the five workload configurations.

Every generator is counter-based (SplitMix64 on seed + index), so a frame is the same on
every machine. Positions are produced in float64 and narrowed to float32 exactly as the
reference's HDF5 loader narrows them (src/main.cpp:315-318).

Workloads (r = particle radius, lattice spacing 2r, jitter U[-0.1r, 0.1r] per axis):
  1  dam-break block [0,.3)x[0,.15)x[0,.3) + droplet sphere R=.05 at (.15,.30,.15), r=.005  (14,015)
  2  sloshing tank x<1, y<.5, z<h(x,y), h=.2+.03 sin(2 pi x) cos(4 pi y), r=.002, + 10k spray (~1.59 M)
  3  crown splash, r=.001: pool [0,.6)^2 x [0,.084) + crown sheet + 5000 log-normal droplets (~4.0 M)
  4  24 sloshing frames, N log-uniform in [50k, 3M], r scaled to N, phase 2 pi k / 24
  5  multi-body pour, r=.0005: pool [0,.5)^2 x [0,.046) + jet R=.02 length .3 + 200k spray (~12.1 M)
"""
import math

import numpy as np

_GOLD = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


class Stream:
    """SplitMix64 counter stream (Steele, Lea & Flood 2014)."""

    def __init__(self, seed):
        self.seed = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
        self.pos = 0

    def u64(self, n):
        with np.errstate(over="ignore"):
            k = np.arange(self.pos + 1, self.pos + 1 + n, dtype=np.uint64)
            z = self.seed + k * _GOLD
            z = (z ^ (z >> np.uint64(30))) * _M1
            z = (z ^ (z >> np.uint64(27))) * _M2
            z = z ^ (z >> np.uint64(31))
        self.pos += n
        return z

    def uniform(self, n, lo=0.0, hi=1.0):
        u = (self.u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * u

    def normal(self, n):
        """Box-Muller; 1-u keeps the log argument in (0, 1]."""
        m = (n + 1) // 2
        u1 = 1.0 - self.uniform(m)
        u2 = self.uniform(m)
        rad = np.sqrt(-2.0 * np.log(u1))
        z = np.concatenate([rad * np.cos(2 * np.pi * u2), rad * np.sin(2 * np.pi * u2)])
        return z[:n]


def _jitter(p, r, st):
    return p + st.uniform(p.size, -0.1 * r, 0.1 * r).reshape(p.shape)


def _lattice(nx, ny, nz, s, origin=(0.0, 0.0, 0.0)):
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    p = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1).astype(np.float64) * s
    return p + np.asarray(origin, dtype=np.float64)


def _ball_offsets(radius, s):
    n = int(math.ceil(radius / s))
    g = np.arange(-n, n + 1, dtype=np.float64) * s
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    o = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
    return o[np.einsum("ij,ij->i", o, o) <= radius * radius]


def dam_break(seed=0x5EED0001):
    """Workload 1: block + droplet sphere on one lattice, r = 0.005."""
    r = 0.005
    s = 2 * r
    block = _lattice(30, 15, 30, s)
    c = np.array([0.15, 0.30, 0.15])
    g = _lattice(31, 41, 31, s)  # covers the sphere's bounding box on the same lattice
    sph = g[np.sum((g - c) ** 2, axis=1) <= 0.05 ** 2 + 1e-12]
    p = np.concatenate([block, sph])
    st = Stream(seed)
    return _jitter(p, r, st).astype(np.float32), r


def _tank_height(x, y, phase=0.0):
    return 0.2 + 0.03 * np.sin(2 * np.pi * x + phase) * np.cos(4 * np.pi * y)


def sloshing_tank(seed=0x5EED0002, r=0.002, phase=0.0, n_spray=10000):
    """Workload 2 (and the frames of workload 4)."""
    s = 2 * r
    nx, ny = int(math.ceil(1.0 / s - 1e-9)), int(math.ceil(0.5 / s - 1e-9))
    xs = np.arange(nx) * s
    ys = np.arange(ny) * s
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    H = _tank_height(X, Y, phase)
    nzc = np.ceil(H / s - 1e-12).astype(np.int64)  # z_k = k s < h
    tot = int(nzc.sum())
    xi = np.repeat(X.ravel(), nzc.ravel())
    yi = np.repeat(Y.ravel(), nzc.ravel())
    starts = np.repeat(np.cumsum(nzc.ravel()) - nzc.ravel(), nzc.ravel())
    zi = (np.arange(tot) - starts) * s
    p = np.stack([xi, yi, zi], axis=1)
    st = Stream(seed)
    # spray: half uniform in the air layer z in [h+4r, 0.35] (rejection), half in 50 clusters
    n_uni = n_spray // 2
    out = []
    while sum(len(a) for a in out) < n_uni:
        q = np.stack([st.uniform(4 * n_uni), st.uniform(4 * n_uni, 0, 0.5), st.uniform(4 * n_uni, 0, 0.35)], 1)
        ok = q[:, 2] >= _tank_height(q[:, 0], q[:, 1], phase) + 4 * r
        out.append(q[ok])
    uni = np.concatenate(out)[:n_uni]
    n_cl = n_spray - n_uni
    ncent = 50
    cent = []
    while len(cent) < ncent:
        q = np.stack([st.uniform(64), st.uniform(64, 0, 0.5), st.uniform(64, 0, 0.35)], 1)
        q = q[q[:, 2] >= _tank_height(q[:, 0], q[:, 1], phase) + 4 * r]
        cent.extend(list(q))
    cent = np.array(cent[:ncent])
    per = np.full(ncent, n_cl // ncent)
    per[: n_cl - per.sum()] += 1
    cl = np.repeat(cent, per, axis=0) + 3 * r * st.normal(3 * n_cl).reshape(-1, 3)
    p = _jitter(p, r, st)
    return np.concatenate([p, uni, cl]).astype(np.float32), r


def crown_splash(seed=0x5EED0003, pool_layers=42, n_drops=5000):
    """Workload 3. pool_layers=42 gives ~4.0 M particles; 20 is the geometry as first stated (~2.03 M)."""
    r = 0.001
    s = 2 * r
    pool = _lattice(300, 300, pool_layers, s)
    top = pool_layers * s
    st = Stream(seed)
    crown = []
    for layer in range(3):
        rho = 0.08 - layer * s
        nth = int(round(2 * np.pi * rho / s))
        th = np.arange(nth) * (2 * np.pi / nth)
        hmax = 0.05 + 0.01 * np.sin(12 * th)
        nz = np.floor(hmax / s).astype(np.int64) + 1
        tt = np.repeat(th, nz)
        starts = np.repeat(np.cumsum(nz) - nz, nz)
        zz = top + (np.arange(nz.sum()) - starts) * s
        crown.append(np.stack([0.3 + rho * np.cos(tt), 0.3 + rho * np.sin(tt), zz], 1))
    crown = np.concatenate(crown)
    radii = 3 * r * np.exp(0.5 * st.normal(n_drops))
    cen = np.stack([st.uniform(n_drops, 0.22, 0.38), st.uniform(n_drops, 0.22, 0.38),
                    st.uniform(n_drops, top + 0.07, top + 0.21)], 1)
    drops = [c + _ball_offsets(rd, s) for c, rd in zip(cen, radii)]
    p = np.concatenate([pool, crown] + drops)
    return _jitter(p, r, st).astype(np.float32), r


def batch_frames(seed=0x5EED0004, n_frames=24):
    """Workload 4: list of (positions, r) for 24 sloshing frames of uneven size."""
    st = Stream(seed)
    ns = np.exp(st.uniform(n_frames, math.log(5e4), math.log(3e6)))
    frames = []
    for k in range(n_frames):
        rk = (0.0125 / ns[k]) ** (1.0 / 3.0)
        nsp = int(round(10000 * ns[k] / 1.578e6))
        frames.append(sloshing_tank(seed + 1000 * (k + 1), r=rk, phase=2 * np.pi * k / n_frames, n_spray=max(nsp, 2)))
    return frames


def multibody_pour(seed=0x5EED0005, pool_layers=46, n_clusters=2000):
    """Workload 5: pool + falling jet + 200k spray, r = 0.0005 (~12.1 M with 46 pool layers)."""
    r = 0.0005
    s = 2 * r
    pool = _lattice(500, 500, pool_layers, s)
    top = pool_layers * s
    st = Stream(seed)
    R = 0.02
    n = int(math.ceil(R / s))
    g = np.arange(-n, n + 1) * s
    X, Y = np.meshgrid(g, g, indexing="ij")
    disk = np.stack([X.ravel(), Y.ravel()], 1)
    disk = disk[np.sum(disk ** 2, 1) <= R * R]
    nz = int(round(0.3 / s))
    jet = np.concatenate([np.column_stack([disk + 0.25, np.full(len(disk), top + k * s)]) for k in range(nz)])
    cen = np.stack([st.uniform(n_clusters, 0, 0.5), st.uniform(n_clusters, 0, 0.5),
                    st.uniform(n_clusters, top + 0.01, top + 0.13)], 1)
    spray = np.repeat(cen, 100, axis=0) + 2 * r * st.normal(3 * 100 * n_clusters).reshape(-1, 3)
    p = _jitter(np.concatenate([pool, jet]), r, st)
    return np.concatenate([p, spray]).astype(np.float32), r


def lattice_sphere(radius=0.05, r=0.005, center=(0.0, 0.0, 0.0), jitter=0.1, seed=7):
    """A ball of lattice particles; used for watertightness checks."""
    p = _ball_offsets(radius, 2 * r) + np.asarray(center)
    st = Stream(seed)
    if jitter:
        p = p + st.uniform(p.size, -jitter * r, jitter * r).reshape(p.shape)
    return p.astype(np.float32), r


def random_cloud(n, r=0.01, extent=0.2, seed=11):
    st = Stream(seed)
    return st.uniform(3 * n, 0.0, extent).reshape(n, 3).astype(np.float32), r


WORKLOADS = {1: dam_break, 2: sloshing_tank, 3: crown_splash, 5: multibody_pour}
