"""Seeded synthetic Al/O slab systems: the inputs SIMOX reads from input.xyz /
restart.xyz (main.f90:205-245), generated instead of read.

The reference ships no input files, so every workload here is built from a
seed: an Al FCC slab (a = 4.05 A, 4-atom basis) with vacuum along z and O atoms
at seeded uniform positions above the top surface.  Coordinates lie in
[-L/2, L/2) as the reference assumes (cellsort.f90:63-69, force.f90:388).
Velocities are Maxwell-Boltzmann in A per TFM fs (main.f90:142), centre-of-mass
motion removed and rescaled per species to the exact temperature, as VINIT
rescales (vverlet.f90:40-50).
"""
from dataclasses import dataclass, field

import numpy as np

AL, NI, FE, O, H = 1, 2, 3, 4, 5
MASS = {AL: 26.981539, NI: 58.6934, FE: 55.84500, O: 15.9994, H: 1.00794}  # datafmt.f90:183,191
KB = 8.617343e-5  # datafmt.f90:27
TFM = 10.180505306645389  # datafmt.f90:24
A_AL = 4.05


@dataclass
class System:
    ids: np.ndarray          # int32 (n,) species 1..5
    x: np.ndarray            # float64 (n, 3)
    v: np.ndarray            # float64 (n, 3), A per TFM fs
    q: np.ndarray            # float64 (n,)
    box: np.ndarray          # float64 (3,)
    fft: tuple               # PME grid (Kx, Ky, Kz)
    rcut: float
    dt: float                # in TFM units (dt_fs / TFM)
    name: str = ""
    meta: dict = field(default_factory=dict)

    @property
    def n(self):
        return int(self.ids.shape[0])


def _place_oxygen(rng, n_o, box, zlo, zhi, dmin):
    """Uniform random O positions in the layer zlo..zhi with a minimum
    separation dmin (periodic in x and y): batched dart throwing with a
    k-d tree; within a batch the earlier candidate wins."""
    from scipy.spatial import cKDTree
    if n_o == 0:
        return np.zeros((0, 3))
    lx, ly = box[0], box[1]
    span = (zhi - zlo) + 4 * dmin  # non-periodic along z: pad the tree's period
    acc = np.zeros((0, 3))
    rounds = 0
    while len(acc) < n_o:
        rounds += 1
        if rounds > 400:
            raise RuntimeError("O placement failed: layer too dense for dmin=%g" % dmin)
        m = max(1024, 2 * (n_o - len(acc)))
        c = np.empty((m, 3))
        c[:, 0] = rng.uniform(0, lx, m)
        c[:, 1] = rng.uniform(0, ly, m)
        c[:, 2] = rng.uniform(0, zhi - zlo, m)
        period = [lx, ly, span]
        ok = np.ones(m, bool)
        if len(acc):
            t = cKDTree(acc, boxsize=period)
            d, _ = t.query(c, k=1, distance_upper_bound=dmin)
            ok &= ~np.isfinite(d)
        c = c[ok]
        if len(c):
            t2 = cKDTree(c, boxsize=period)
            drop = np.zeros(len(c), bool)
            for i, j in sorted(t2.query_pairs(dmin)):
                if not drop[i]:
                    drop[j] = True
            c = c[~drop]
        acc = np.concatenate([acc, c[: n_o - len(acc)]], 0)
    out = acc.copy()
    out[:, 0] -= lx / 2
    out[:, 1] -= ly / 2
    out[:, 2] += zlo
    return out


def maxwell_boltzmann(rng, ids, T):
    """v ~ N(0, sqrt(kB T / m)) per component, COM momentum removed, then each
    species rescaled to exactly T (vverlet.f90:40-50 rescales per species)."""
    n = ids.shape[0]
    m = np.array([MASS[int(i)] for i in ids])
    kT = KB * T
    v = rng.normal(size=(n, 3)) * np.sqrt(kT / m)[:, None]
    p = (m[:, None] * v).sum(0) / m.sum()
    v -= p
    for s in np.unique(ids):
        sel = ids == s
        if sel.sum() == 0 or T <= 0:
            continue
        mv2 = (m[sel, None] * v[sel] ** 2).sum()
        T0 = mv2 / (3 * sel.sum())
        v[sel] *= np.sqrt(kT / T0)
    return v


def al_o_slab(nx, ny, nz, n_o, seed, T, fft, rcut, vacuum=20.0, o_layer=(2.0, 6.0),
              o_dmin=2.0, dt_fs=1.0, name=""):
    rng = np.random.default_rng(seed)
    a = A_AL
    box = np.array([nx * a, ny * a, nz * a + vacuum])
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]]) * a
    gi, gj, gk = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    cells = np.stack([gi.ravel(), gj.ravel(), gk.ravel()], 1) * a
    al = (cells[:, None, :] + basis[None, :, :]).reshape(-1, 3)
    zlo = -box[2] / 2 + vacuum / 2
    al[:, 0] -= box[0] / 2
    al[:, 1] -= box[1] / 2
    al[:, 2] += zlo
    z_top = al[:, 2].max()
    ox = _place_oxygen(rng, n_o, box, z_top + o_layer[0], z_top + o_layer[1], o_dmin)
    x = np.concatenate([al, ox], 0)
    ids = np.concatenate([np.full(len(al), AL, np.int32), np.full(n_o, O, np.int32)])
    # keep coordinates inside [-L/2, L/2)
    x = x - box * np.floor((x + box / 2) / box)
    v = maxwell_boltzmann(rng, ids, T)
    q = np.zeros(len(ids))
    return System(ids=ids, x=x, v=v, q=q, box=box, fft=tuple(fft), rcut=float(rcut),
                  dt=dt_fs / TFM, name=name,
                  meta=dict(nx=nx, ny=ny, nz=nz, n_o=n_o, seed=seed, T=T, vacuum=vacuum))


def config(name, seed=1):
    """The named workloads.  C1: 6x6x8 Al (1152) + 64 O, 300 K, grid 32x32x96.
    C2: 24x24x32 Al (73,728) + 4,096 O, 600 K, grid 96x96x192.
    C3: 64x64x60 Al (983,040) + 65,536 O, 900 K, grid 256x256x512.
    c3_sample: a sixteenth of C3 by the same recipe (16x16x60 Al + 4,096 O,
    grid 64x64x512: the same atoms and grid points per unit of surface), for
    the CPU timings."""
    if name == "c1":
        return al_o_slab(6, 6, 8, 64, seed, 300.0, (32, 32, 96), 8.0, vacuum=20.0,
                         o_layer=(2.0, 6.0), o_dmin=2.0, name="c1")
    if name == "c1_rcut10":
        return al_o_slab(6, 6, 8, 64, seed, 300.0, (32, 32, 96), 10.0, vacuum=20.0,
                         o_layer=(2.0, 6.0), o_dmin=2.0, name="c1_rcut10")
    if name == "c2":
        return al_o_slab(24, 24, 32, 4096, seed, 600.0, (96, 96, 192), 10.0, vacuum=20.0,
                         o_layer=(2.0, 10.0), o_dmin=2.0, name="c2")
    if name == "c3":
        return al_o_slab(64, 64, 60, 65536, seed, 900.0, (256, 256, 512), 10.0, vacuum=28.0,
                         o_layer=(2.0, 10.0), o_dmin=1.5, name="c3")
    if name == "c3_sample":
        return al_o_slab(16, 16, 60, 4096, seed, 900.0, (64, 64, 512), 10.0, vacuum=28.0,
                         o_layer=(2.0, 10.0), o_dmin=1.5, name="c3_sample")
    raise KeyError(name)


def fft_size(n):
    """Smallest K >= n that is a multiple of 4 with no prime factor above 5."""
    k = max(4, int(np.ceil(n / 4.0)) * 4)
    while True:
        m = k
        for p in (2, 3, 5):
            while m % p == 0:
                m //= p
        if m == 1:
            return k
        k += 4


def replicas(n_rep=256, seed=1):
    """Independent Al/O slabs of uneven size: slab cells drawn from 4x4x6 to
    10x10x12 (384-4,800 Al), O coverage 0.1-1.0 monolayer of the (001) surface
    (2 sites per a^2), temperature 300-1200 K.  Rcut = min(10, L_xy / 3) keeps
    3 or more cells per dimension; PME grid spacing is Rcut / 10 per dimension."""
    rng = np.random.default_rng(seed)
    out = []
    for r in range(n_rep):
        nx, ny = (int(v) for v in rng.integers(4, 11, size=2))
        nz = int(rng.integers(6, 13))
        cov = float(rng.uniform(0.1, 1.0))
        T = float(rng.uniform(300.0, 1200.0))
        n_o = max(1, int(round(cov * 2 * nx * ny)))
        lxy = min(nx, ny) * A_AL
        rcut = min(10.0, np.floor(lxy / 3.0 * 100.0) / 100.0)
        vacuum = 20.0
        box = np.array([nx * A_AL, ny * A_AL, nz * A_AL + vacuum])
        fft = tuple(fft_size(b / (rcut / 10.0)) for b in box)
        out.append(al_o_slab(nx, ny, nz, n_o, seed * 100003 + r, T, fft, rcut, vacuum=vacuum,
                             o_layer=(2.0, 6.0), o_dmin=2.0, name="c4_%d" % r))
    return out


def tiny(seed=1, nx=4, ny=4, nz=4, n_o=16, T=300.0, fft=(16, 16, 32), rcut=6.0, vacuum=12.0):
    """A small system for fast parity tests."""
    return al_o_slab(nx, ny, nz, n_o, seed, T, fft, rcut, vacuum=vacuum, o_layer=(2.0, 5.0),
                     o_dmin=2.0, name="tiny")


def tiny_alloy(seed=1, nx=5, ny=5, nz=4, n_o=16, frac_ni=0.15, frac_fe=0.10, fft=(20, 20, 32),
               rcut=6.0):
    """A small Al slab with seeded Ni (id 2) and Fe (id 3) substitutions and an
    O layer: exercises the Al-Ni (AL_NI) and other metal-metal (METAL_HETERO)
    pair branches.  3 or more cells per dimension."""
    s = al_o_slab(nx, ny, nz, n_o, seed, 300.0, fft, rcut, vacuum=12.0, o_layer=(2.0, 5.0),
                  o_dmin=2.0, name="tiny_alloy")
    rng = np.random.default_rng(seed + 7919)
    ids = s.ids.copy()
    metal = np.flatnonzero(ids == AL)
    pick = rng.permutation(metal)
    n_ni, n_fe = int(round(frac_ni * len(metal))), int(round(frac_fe * len(metal)))
    ids[pick[:n_ni]] = NI
    ids[pick[n_ni:n_ni + n_fe]] = FE
    v = maxwell_boltzmann(rng, ids, 300.0)
    return System(ids=ids, x=s.x, v=v, q=s.q, box=s.box, fft=s.fft, rcut=s.rcut, dt=s.dt,
                  name="tiny_alloy", meta=dict(s.meta, frac_ni=frac_ni, frac_fe=frac_fe))


def snapshots(n_snap=64, seed=5):
    """Seeded snapshots for parameter fitting: Al slab 10x10x12 cells (4,800 Al)
    + 400 O in a 2-10 A layer, temperatures 300-1200 K, Rcut 10 A (4 x 4 x 6
    cells), PME grid 40 x 40 x 72."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_snap):
        T = float(rng.uniform(300.0, 1200.0))
        out.append(al_o_slab(10, 10, 12, 400, seed * 1000 + k, T, (40, 40, 72), 10.0,
                             vacuum=20.0, o_layer=(2.0, 10.0), o_dmin=2.0, name="c5_%d" % k))
    return out

