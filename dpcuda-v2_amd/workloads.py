"""Synthetic inputs for the benchmark configurations (numpy only, no device work).

Bidispersity follows Particle::setBiDispersity (src/particles/base/particle.cu:926-944):
the first floor(N * count_ratio) particles are large (diameter size_ratio), the rest
have diameter 1.  The box side for a packing fraction follows
Particle::scaleToPackingFraction (particle.cu:1037-1045): L = sqrt(pi * sum r^2 / phi).
Positions come from numpy's seeded generator (the reference's Thrust stream is not
needed: parity runs feed the same arrays to the GPU and to the checker).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

PI = 3.141592653589793238462643383279502884197


def bidisperse_radii(n: int, size_ratio: float = 1.4, count_ratio: float = 0.5) -> np.ndarray:
    n_large = int(n * count_ratio)
    r = np.full(n, 0.5)
    r[:n_large] = size_ratio / 2.0
    return r


def box_side(radii: np.ndarray, phi: float) -> float:
    return math.sqrt(float(np.sum(radii * radii)) * PI / phi)


def uniform_positions(n: int, side: float, seed: int) -> tuple[np.ndarray, np.ndarray]:
    rng = np.random.default_rng(seed)
    return rng.uniform(0.0, side, n), rng.uniform(0.0, side, n)


def gaussian_velocities(n: int, temperature: float, seed: int, n_dim: int = 2):
    """Normal draws with mean removed and rescaled to exactly T
    (Particle::setRandomVelocities, particle.cu:908-912), done on the host."""
    rng = np.random.default_rng(seed + 7919)
    vx = rng.normal(0.0, math.sqrt(temperature), n)
    vy = rng.normal(0.0, math.sqrt(temperature), n)
    vx -= vx.mean()
    vy -= vy.mean()
    t = (0.5 * (vx * vx + vy * vy)).sum() * 2.0 / (n_dim * n)
    f = math.sqrt(temperature / t)
    return vx * f, vy * f


@dataclass
class NeighborSpec:
    """Neighbour-list numbers for "cell size = max cutoff + skin" and "rebuild when
    max displacement > skin/2" (the reference triggers on 3*max d^2 > thr^2,
    particle.cu:1141-1161)."""
    method: int
    cutoff: float
    disp_thresh_sq: float
    n_cells_dim: int
    cell_disp_thresh_sq: float


def neighbor_spec(side: float, max_diam: float = 1.4, skin: float = 0.3,
                  method: int = 3) -> NeighborSpec:
    rc = max_diam + skin
    half_skin = skin / 2.0
    thr_sq = 3.0 * half_skin * half_skin
    ncd = max(3, int(math.floor(side / rc)))
    cs = side / ncd
    cell_thr_sq = (0.5 * cs) ** 2
    return NeighborSpec(method, rc, thr_sq, ncd, cell_thr_sq)


def configure(sys_, spec: NeighborSpec):
    sys_.neighbor_config(spec.method, spec.cutoff, spec.disp_thresh_sq, spec.n_cells_dim,
                         spec.cell_disp_thresh_sq)


@dataclass
class DiskState:
    n: int
    side: float
    x: np.ndarray
    y: np.ndarray
    rad: np.ndarray
    mass: np.ndarray


def disk_config(n: int, phi: float, seed: int = 12345) -> DiskState:
    rad = bidisperse_radii(n)
    side = box_side(rad, phi)
    x, y = uniform_positions(n, side, seed)
    return DiskState(n, side, x, y, rad, np.ones(n))


def make_disk_system(sys_cls, st: DiskState, spec: NeighborSpec | None = None, backend=None,
                     e_c: float = 1.0, n_c: float = 2.0):
    """Create, fill and neighbour-initialise a disk system on `backend`."""
    s = sys_cls(0, st.n, 0, backend=backend)
    s.set_box(st.side, st.side)
    s.set_energy(e_c, n_c)
    s.set_state(st.x, st.y, st.rad, st.mass)
    configure(s, spec or neighbor_spec(st.side))
    s.neighbor_init()
    return s


# rigid bumpy geometry: RigidBumpy::calculateVertexRadius (rigid_bumpy.cu:570-579)
def vertex_radius(d_min: float = 1.0, n_small: int = 20, seg: float = 1.0) -> float:
    ang = 2 * math.pi / n_small
    return d_min / (1 + seg / math.sin(ang / 2)) / 2.0


@dataclass
class BumpyState:
    n: int
    nv: int
    side: float
    x: np.ndarray
    y: np.ndarray
    angle: np.ndarray
    rad: np.ndarray
    mass: np.ndarray
    nverts: np.ndarray
    rv: float


def bumpy_config(n: int, phi: float, seed: int = 12345, n_small: int = 20,
                 n_large: int = 28) -> BumpyState:
    rad = bidisperse_radii(n)
    nverts = np.where(rad == 0.7, n_large, n_small).astype(np.int64)
    side = box_side(rad, phi)
    x, y = uniform_positions(n, side, seed)
    rng = np.random.default_rng(seed + 1)
    angle = rng.uniform(0.0, 2 * math.pi, n)
    return BumpyState(n, int(nverts.sum()), side, x, y, angle, rad, np.ones(n), nverts,
                      vertex_radius(1.0, n_small, 1.0))


# RigidBumpy::initializeFromConfig scales the configured contact energy by the square
# of the geometry scale 2 r_v / d_min (rigid_bumpy.cu:237-241, :1265-1269)
def bumpy_contact_scale(st: BumpyState, e_c: float = 1.0) -> float:
    return e_c * (2.0 * st.rv / (2.0 * float(st.rad.min()))) ** 2


def make_bumpy_system(sys_cls, st: BumpyState, backend=None, e_c: float = 1.0,
                      rotation: bool = True, method: int = 3):
    s = sys_cls(1, st.n, st.nv, backend=backend)
    s.set_box(st.side, st.side)
    s.set_energy(e_c, 2.0)
    s.set_state(st.x, st.y, st.rad, st.mass)
    s.upload(10, st.angle)
    s.bumpy_setup(st.nverts, st.rv, rotation)
    # RigidBumpy::setNeighborSize (rigid_bumpy.cu:1278-1300) with multiplier 1.5 /
    # displacement multiplier 0.2 and segment ratio 1 (standard_configs.h:58-65)
    rc = 1.5 * 1.4
    spec = neighbor_spec(st.side, 1.4, rc - 1.4, method)
    thr_sq = (0.2 * 1.0 * st.rv) ** 2
    s.neighbor_config(method, rc, thr_sq, spec.n_cells_dim, spec.cell_disp_thresh_sq)
    s.bumpy_neighbor_config(1.5 * 2.0 * st.rv, 1.4)
    s.neighbor_init()
    return s


# ---- batched and served jamming workloads -------------------------------------------
def batched_sizes(n_systems: int, seed: int, n_lo: int = 64, n_hi: int = 2048) -> np.ndarray:
    """Particles per system, log-uniform in [n_lo, n_hi] from the seed."""
    rng = np.random.default_rng(seed)
    return np.exp(rng.uniform(np.log(n_lo), np.log(n_hi), n_systems)).astype(np.int64)


def batched_systems(n_systems: int, seed: int, n_lo: int = 64, n_hi: int = 2048,
                    phi0: float = 0.80) -> list:
    """Independent bidisperse disk systems of uneven N at packing fraction phi0, each
    with its own seeded uniform positions."""
    ns = batched_sizes(n_systems, seed, n_lo, n_hi)
    return [disk_config(int(n), phi0, seed=seed + 1 + k) for k, n in enumerate(ns)]


@dataclass
class ServeRequest:
    arrival: float    # seconds after the start
    bumpy: bool
    state: object     # DiskState or BumpyState
    pressure: float   # target pressure


def serve_requests(count: int, rate: float, seed: int, bumpy_fraction: float = 0.3,
                   disk_n=(128, 16384), bumpy_n=(128, 4096), pressure=(1e-7, 1e-3),
                   phi0: float = 0.80) -> list:
    """Poisson arrivals at `rate` per second: a fraction `bumpy_fraction` rigid bumpy
    systems (20/28 vertices, N log-uniform in bumpy_n, at phi0 - 0.1), the rest disks (N
    log-uniform in disk_n, at phi0); target pressure log-uniform in `pressure`."""
    rng = np.random.default_rng(seed)
    out, t = [], 0.0
    for k in range(count):
        t += rng.exponential(1.0 / rate)
        bumpy = bool(rng.uniform() < bumpy_fraction)
        pt = float(np.exp(rng.uniform(np.log(pressure[0]), np.log(pressure[1]))))
        lo, hi = bumpy_n if bumpy else disk_n
        n = int(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        st = (bumpy_config(n, phi0 - 0.1, seed=seed + 1000 + k) if bumpy
              else disk_config(n, phi0, seed=seed + 1000 + k))
        out.append(ServeRequest(t, bumpy, st, pt))
    return out
