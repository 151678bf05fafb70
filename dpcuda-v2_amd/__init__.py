"""dpcuda-v2_amd -- MI355X-native particle hot path of dpCUDA-v2 (Python host side).

The compute lives in ``libdpm.so`` (hand-written gfx950 HIP kernels behind the C
ABI in ``include/dpm.h``).  This module is a thin ctypes layer over that ABI:
``System`` mirrors the reference's ``Particle`` operations
(include/particles/base/particle.h:28-717) one call per entry point, moving
numpy arrays in static particle order.

There is no CPU fallback: if ``libdpm.so`` is missing or no HIP device is
visible, ``load()`` / ``System()`` raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdpm.so")

DISK, BUMPY = 0, 1
NBR_NONE, NBR_ALL, NBR_VERLET, NBR_CELL = 0, 1, 2, 3

F_X, F_Y, F_VX, F_VY, F_FX, F_FY, F_RAD, F_MASS, F_PE, F_KE = range(10)
F_ANGLE, F_OMEGA, F_TORQUE, F_MOI = 10, 11, 12, 13
F_STATIC_INDEX, F_NUM_NEIGHBORS, F_CONTACTS = 14, 15, 16
F_STRESS_XX, F_STRESS_XY, F_STRESS_YX, F_STRESS_YY = 17, 18, 19, 20
F_VERTEX_X, F_VERTEX_Y, F_VERTEX_MASS, F_NUM_VERTICES, F_VERTEX_CONTACTS = 21, 22, 23, 24, 25
INT_FIELDS = {F_STATIC_INDEX, F_NUM_NEIGHBORS, F_CONTACTS, F_NUM_VERTICES, F_VERTEX_CONTACTS}
VERTEX_FIELDS = {F_VERTEX_X, F_VERTEX_Y, F_VERTEX_MASS}
# pair analytics fields (dpm.h DPM_P_*), in field order; the first 18 exist for disks
PAIR_FIELDS = ("potential_pairs", "force_pairs_x", "force_pairs_y", "distance_pairs_x",
               "distance_pairs_y", "pair_ids_i", "pair_ids_j", "overlap_pairs", "radsum_pairs",
               "pair_separation_angle", "hessian_pairs_xx", "hessian_pairs_xy",
               "hessian_pairs_yx", "hessian_pairs_yy", "hessian_ii_xx", "hessian_ii_xy",
               "hessian_ii_yx", "hessian_ii_yy", "angle_pairs_i", "angle_pairs_j",
               "this_vertex_contact_counts", "pair_friction_coefficient", "pair_vertex_overlaps",
               "hessian_pairs_xt", "hessian_pairs_yt", "hessian_pairs_tx", "hessian_pairs_ty",
               "hessian_pairs_tt", "hessian_ii_xt", "hessian_ii_yt", "hessian_ii_tx",
               "hessian_ii_ty", "hessian_ii_tt")
PAIR_DISK_FIELDS = 18
PAIR_INT_FIELDS = {5, 6, 20}

(RED_PE, RED_KE, RED_POWER, RED_FMAX_SQ, RED_FORCE_BALANCE, RED_PRESSURE, RED_CONTACTS,
 RED_TEMPERATURE, RED_PACKING_FRACTION, RED_MAX_NEIGH_DISP_SQ, RED_MAX_CELL_DISP_SQ,
 RED_STRESS_XX, RED_STRESS_XY, RED_STRESS_YX, RED_STRESS_YY, RED_VERTEX_CONTACTS) = range(16)

OPT_GRAPH, OPT_CHUNK, OPT_TIMING = 0, 1, 2
OPT_FORCE_BLOCK, OPT_CELL_PCT = 3, 4
K_FORCE, K_DRIFT, K_REBUILD = 0, 1, 2


class DpmError(RuntimeError):
    pass


class FireParams(C.Structure):
    _fields_ = [("alpha_init", C.c_double), ("dt", C.c_double),
                ("avg_pe_target", C.c_double), ("avg_pe_diff_target", C.c_double),
                ("fmax_target", C.c_double), ("max_steps", C.c_int64)]


class FireResult(C.Structure):
    _fields_ = [("steps", C.c_int64), ("stopped", C.c_int32), ("converged", C.c_int32),
                ("avg_pe", C.c_double), ("fmax_sq", C.c_double), ("dt", C.c_double),
                ("alpha", C.c_double)]


class FireState(C.Structure):
    _fields_ = [("alpha", C.c_double), ("alpha_init", C.c_double), ("dt", C.c_double),
                ("dt_max", C.c_double), ("dt_min", C.c_double), ("n_good", C.c_int64),
                ("n_bad", C.c_int64), ("stopped", C.c_int32), ("power", C.c_double)]


class JamParams(C.Structure):
    _fields_ = [("pressure_target", C.c_double), ("window_rel", C.c_double),
                ("strain_step", C.c_double), ("decompress_factor", C.c_double),
                ("length_tol", C.c_double), ("min_contacts_per_particle", C.c_double),
                ("max_compression_steps", C.c_int64), ("fire", FireParams)]


class JamResult(C.Structure):
    _fields_ = [("compression_steps", C.c_int64), ("fire_steps_total", C.c_int64),
                ("status", C.c_int32), ("pressure", C.c_double),
                ("packing_fraction", C.c_double), ("pe", C.c_double),
                ("box_length", C.c_double), ("contacts", C.c_int64)]


def fire_params(alpha_init=0.1, dt=1e-2, avg_pe_target=1e-16, avg_pe_diff_target=1e-16,
                fmax_target=0.0, max_steps=100000):
    """Defaults of minimizeFire(Particle&) (src/routines/minimization.cu:142-148)."""
    return FireParams(alpha_init, dt, avg_pe_target, avg_pe_diff_target, fmax_target,
                      int(max_steps))


class Request(C.Structure):
    _fields_ = [("kind", C.c_int32), ("mode", C.c_int32), ("n_particles", C.c_int64),
                ("n_vertices", C.c_int64), ("x", C.c_void_p), ("y", C.c_void_p),
                ("radius", C.c_void_p), ("mass", C.c_void_p), ("angle", C.c_void_p),
                ("vertices_per_particle", C.c_void_p), ("vertex_radius", C.c_double),
                ("box_side", C.c_double), ("e_c", C.c_double), ("jam", JamParams)]


def jam_params(pressure_target=1e-7, window_rel=1e-2, strain_step=1e-4,
               decompress_factor=1.01, length_tol=1e-10, min_contacts_per_particle=0.0,
               max_compression_steps=100000, fire=None):
    return JamParams(pressure_target, window_rel, strain_step, decompress_factor, length_tol,
                     min_contacts_per_particle, int(max_compression_steps),
                     fire if fire is not None else fire_params())


_P = C.c_void_p
_D = C.c_double
_I = C.c_int
_L = C.c_int64
_SIGS = {
    "create": (_P, [_I, _L, _L]),
    "destroy": (None, [_P]),
    "last_error": (C.c_char_p, []),
    "set_box": (_I, [_P, _D, _D]),
    "get_box": (_I, [_P, C.POINTER(_D), C.POINTER(_D)]),
    "set_energy": (_I, [_P, _D, _D]),
    "upload": (_I, [_P, _I, _P, _L]),
    "download": (_I, [_P, _I, _P, _L]),
    "neighbor_config": (_I, [_P, _I, _D, _D, _L, _D]),
    "bumpy_neighbor_config": (_I, [_P, _D, _D]),
    "neighbor_init": (_I, [_P]),
    "neighbor_check": (_I, [_P, C.POINTER(_I)]),
    "neighbor_pairs": (_I, [_P, _P, _P, _L, C.POINTER(_L)]),
    "vertex_pairs": (_I, [_P, _P, _P, _L, C.POINTER(_L)]),
    "neighbor_stats": (_I, [_P, C.POINTER(_L), C.POINTER(_L), C.POINTER(_L), C.POINTER(_L)]),
    "zero_forces": (_I, [_P]),
    "compute_forces": (_I, [_P]),
    "kick": (_I, [_P, _D]),
    "drift": (_I, [_P, _D]),
    "mix": (_I, [_P, _D]),
    "zero_velocities": (_I, [_P]),
    "nve_run": (_I, [_P, _D, _L]),
    "reduce": (_I, [_P, _I, C.POINTER(_D)]),
    "minimize_fire": (_I, [_P, C.POINTER(FireParams), C.POINTER(FireResult)]),
    "fire_state_init": (_I, [C.POINTER(FireState), _D, _D]),
    "fire_step": (_I, [_P, C.POINTER(FireState)]),
    "adam_reset": (_I, [_P]),
    "adam_step": (_I, [_P, _L, _D, _D, _D, _D]),
    "minimize_adam": (_I, [_P, _D, _D, _D, _D, _D, _D, _L, C.POINTER(_L)]),
    "compute_stress": (_I, [_P]),
    "count_contacts": (_I, [_P]),
    "scale_positions": (_I, [_P, _D]),
    "scale_box": (_I, [_P, _D]),
    "scale_to_packing_fraction": (_I, [_P, _D]),
    "snapshot_save": (_I, [_P]),
    "snapshot_restore": (_I, [_P]),
    "remove_mean_velocity": (_I, [_P]),
    "scale_to_temperature": (_I, [_P, _D]),
    "jam": (_I, [_P, C.POINTER(JamParams), C.POINTER(JamResult)]),
    "bumpy_setup": (_I, [_P, _P, _D, _I]),
    "pair_analytics": (_I, [_P, C.POINTER(_L)]),
    "pair_download": (_I, [_P, _I, _P, _L]),
    "bumpy_area": (_I, [_P, _P, _L]),
}
# entry points only the HIP library has
_GPU_ONLY = {
    "device_count": (_I, []),
    "set_option": (_I, [_P, _I, _L]),
    "kernel_time": (_I, [_P, _I, C.POINTER(_D), C.POINTER(_L)]),
    "reset_kernel_time": (_I, [_P]),
    "synchronize": (_I, [_P]),
    "batch_create": (_P, [_L, _P]),
    "batch_create_mixed": (_P, [_L, _P, _P]),
    "batch_set_rigid_system": (_I, [_P, _L, _P, _P, _P, _P, _P, _P, _D, _D, C.c_int32, _D]),
    "batch_get_rigid_system": (_I, [_P, _L, _P, _P, _P, _P, _P, C.POINTER(_D)]),
    "batch_destroy": (None, [_P]),
    "batch_set_system": (_I, [_P, _L, _P, _P, _P, _P, _D]),
    "batch_get_system": (_I, [_P, _L, _P, _P, C.POINTER(_D)]),
    "batch_config": (_I, [_P, _D, _D, _D, _D]),
    "batch_set_option": (_I, [_P, _I, _L]),
    "batch_jam": (_I, [_P, C.POINTER(JamParams), _L, _P, C.POINTER(_L)]),
    "batch_finish_steps": (_I, [_P, _P]),
    "batch_minimize": (_I, [_P, C.POINTER(FireParams), _L, _P, C.POINTER(_L)]),
    "batch_kernel_time": (_I, [_P, C.POINTER(_D), C.POINTER(_L)]),
    "batch_row_entries": (_I, [_P, C.POINTER(_L)]),
    "batch_vertex_row_entries": (_I, [_P, C.POINTER(_L)]),
    "server_release": (_I, [_P, _L]),
    "disk_forces_noalloc": (_I, [_P, _P, _P, _L, _D, _D, _D, _D, _D, _P, _P, _P, _P, _P]),
    "server_create": (_P, [_L, _L, _L, _L]),
    "server_destroy": (None, [_P]),
    "server_submit": (_L, [_P, C.POINTER(Request)]),
    "server_progress": (_I, [_P, _L, C.POINTER(_L), C.POINTER(_L), C.POINTER(_L)]),
    "server_result": (_I, [_P, _L, C.POINTER(C.c_int32), C.POINTER(JamResult), C.POINTER(_D),
                           _P, _P, C.POINTER(_D)]),
    "batch_autotune": (_I, [_P, C.POINTER(JamParams), _L, _P, _I, _P, _I, C.POINTER(_D),
                           C.POINTER(_L), C.POINTER(_D)]),
    "batch_autotune_shape": (_I, [_P, C.POINTER(JamParams), _L, C.POINTER(_L), C.POINTER(_L),
                                  C.POINTER(_D)]),
}


class Backend:
    """One C library implementing the dpm_ call set (or the oracle's orc_ mirror)."""

    def __init__(self, path: str, prefix: str, gpu: bool):
        if not os.path.exists(path):
            raise DpmError(f"{path} is not built (run __graft_entry__.build())")
        self.path = path
        self.prefix = prefix
        self.gpu = gpu
        self.lib = C.CDLL(path)
        sigs = dict(_SIGS)
        if gpu:
            sigs.update(_GPU_ONLY)
        self.fn = {}
        for name, (res, args) in sigs.items():
            f = getattr(self.lib, prefix + name)
            f.restype = res
            f.argtypes = args
            self.fn[name] = f

    def err(self) -> str:
        m = self.fn["last_error"]()
        return m.decode() if m else ""

    def call(self, name, *args):
        r = self.fn[name](*args)
        if r is not None and isinstance(r, int) and r < 0:
            raise DpmError(f"{self.prefix}{name}: {self.err()}")
        return r


_gpu_backend = None


def load(path: str = LIB_PATH) -> Backend:
    """Load libdpm.so.  Raises DpmError when it is missing (there is no fallback)."""
    global _gpu_backend
    if _gpu_backend is None or _gpu_backend.path != path:
        _gpu_backend = Backend(path, "dpm_", gpu=True)
    return _gpu_backend


def device_count() -> int:
    return int(load().fn["device_count"]())


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class System:
    """One particle system on one GPU (or, for tests, on the CPU oracle backend)."""

    def __init__(self, kind: int, n_particles: int, n_vertices: int = 0,
                 backend: Backend | None = None):
        self.be = backend if backend is not None else load()
        if self.be.gpu and self.be.fn["device_count"]() <= 0:
            raise DpmError("no HIP device visible: the HIP path cannot run")
        self.kind = kind
        self.n = int(n_particles)
        self.nv = int(n_vertices)
        h = self.be.fn["create"](kind, self.n, self.nv)
        if not h:
            raise DpmError(f"{self.be.prefix}create: {self.be.err()}")
        self.h = C.c_void_p(h)

    def close(self):
        if getattr(self, "h", None):
            self.be.fn["destroy"](self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _c(self, name, *args):
        return self.be.call(name, self.h, *args)

    # parameters
    def set_box(self, lx, ly=None):
        self._c("set_box", float(lx), float(lx if ly is None else ly))

    def box(self):
        a, b = C.c_double(), C.c_double()
        self._c("get_box", C.byref(a), C.byref(b))
        return a.value, b.value

    def set_energy(self, e_c=1.0, n_c=2.0):
        self._c("set_energy", float(e_c), float(n_c))

    # data
    def upload(self, field: int, values):
        if field == F_NUM_VERTICES:
            a = np.ascontiguousarray(values, dtype=np.int64)
        else:
            a = np.ascontiguousarray(values, dtype=np.float64)
        self._c("upload", field, _ptr(a), a.size)

    def download(self, field: int) -> np.ndarray:
        n = self.nv if field in VERTEX_FIELDS else self.n
        a = np.empty(n, dtype=np.int64 if field in INT_FIELDS else np.float64)
        self._c("download", field, _ptr(a), n)
        return a

    def set_state(self, x, y, rad, mass, vx=None, vy=None):
        self.upload(F_X, x)
        self.upload(F_Y, y)
        self.upload(F_RAD, rad)
        self.upload(F_MASS, mass)
        self.upload(F_VX, np.zeros(self.n) if vx is None else vx)
        self.upload(F_VY, np.zeros(self.n) if vy is None else vy)

    # neighbours
    def neighbor_config(self, method, cutoff=0.0, disp_thresh_sq=0.0, n_cells_dim=0,
                        cell_disp_thresh_sq=0.0):
        self._c("neighbor_config", int(method), float(cutoff), float(disp_thresh_sq),
                int(n_cells_dim), float(cell_disp_thresh_sq))

    def bumpy_neighbor_config(self, vertex_cutoff, vertex_particle_cutoff):
        self._c("bumpy_neighbor_config", float(vertex_cutoff), float(vertex_particle_cutoff))

    def neighbor_init(self):
        self._c("neighbor_init")

    def neighbor_check(self) -> int:
        r = C.c_int(0)
        self._c("neighbor_check", C.byref(r))
        return r.value

    def _pairs(self, name, cap):
        cnt = C.c_int64(0)
        while True:
            oi = np.empty(cap, np.int64)
            oj = np.empty(cap, np.int64)
            r = self.be.fn[name](self.h, _ptr(oi), _ptr(oj), cap, C.byref(cnt))
            if r == 0:
                return oi[: cnt.value], oj[: cnt.value]
            if cnt.value > cap:
                cap = int(cnt.value)
                continue
            raise DpmError(f"{self.be.prefix}{name}: {self.be.err()}")

    def neighbor_pairs(self):
        return self._pairs("neighbor_pairs", max(16, 16 * self.n))

    def vertex_pairs(self):
        return self._pairs("vertex_pairs", max(16, 16 * self.nv))

    def pair_analytics(self) -> dict:
        """Per-slot pair arrays in the reference's layout: {name: (N, row) array}; rows are
        the library's internal particle order, unused slots hold the fill values."""
        row = C.c_int64(0)
        self._c("pair_analytics", C.byref(row))
        out = {}
        nf = len(PAIR_FIELDS) if self.kind == BUMPY else PAIR_DISK_FIELDS
        for f, name in enumerate(PAIR_FIELDS[:nf]):
            a = np.empty(self.n * row.value, np.int64 if f in PAIR_INT_FIELDS else np.float64)
            self._c("pair_download", f, _ptr(a), a.size)
            out[name] = a.reshape(self.n, row.value)
        return out

    def bumpy_area(self) -> np.ndarray:
        a = np.empty(self.n)
        self._c("bumpy_area", _ptr(a), self.n)
        return a

    def neighbor_stats(self, totals=True):
        """totals False skips the row-entry total (a download of every row length)."""
        v = [C.c_int64() for _ in range(4)]
        self._c("neighbor_stats", *[C.byref(x) if (k != 1 or totals) else None
                                    for k, x in enumerate(v)])
        return dict(max_alloc=v[0].value, total_entries=v[1].value, n_rebuilds=v[2].value,
                    n_cell_rebuilds=v[3].value)

    # dynamics
    def zero_forces(self):
        self._c("zero_forces")

    def compute_forces(self):
        self._c("compute_forces")

    def kick(self, dt):
        self._c("kick", float(dt))

    def drift(self, dt):
        self._c("drift", float(dt))

    def mix(self, alpha):
        self._c("mix", float(alpha))

    def zero_velocities(self):
        self._c("zero_velocities")

    def nve_run(self, dt, n_steps):
        self._c("nve_run", float(dt), int(n_steps))

    def reduce(self, what: int) -> float:
        r = C.c_double()
        self._c("reduce", int(what), C.byref(r))
        return r.value

    def minimize_fire(self, params: FireParams | None = None) -> FireResult:
        res = FireResult()
        self._c("minimize_fire", C.byref(params or fire_params()), C.byref(res))
        return res

    def fire_state(self, alpha_init=0.1, dt=1e-2) -> FireState:
        st = FireState()
        self.be.call("fire_state_init", C.byref(st), float(alpha_init), float(dt))
        return st

    def fire_step(self, st: FireState):
        self._c("fire_step", C.byref(st))

    def adam_reset(self):
        self._c("adam_reset")

    def adam_step(self, step, alpha, beta1=0.9, beta2=0.999, eps=1e-8):
        self._c("adam_step", int(step), float(alpha), float(beta1), float(beta2), float(eps))

    def minimize_adam(self, alpha=1e-4, beta1=0.9, beta2=0.999, eps=1e-8, avg_pe_target=1e-16,
                      avg_pe_diff_target=1e-16, max_steps=100000) -> int:
        st = C.c_int64()
        self._c("minimize_adam", float(alpha), float(beta1), float(beta2), float(eps),
                float(avg_pe_target), float(avg_pe_diff_target), int(max_steps), C.byref(st))
        return st.value

    def compute_stress(self):
        self._c("compute_stress")

    def count_contacts(self):
        self._c("count_contacts")

    def scale_positions(self, f):
        self._c("scale_positions", float(f))

    def scale_box(self, f):
        self._c("scale_box", float(f))

    def scale_to_packing_fraction(self, phi):
        self._c("scale_to_packing_fraction", float(phi))

    def snapshot_save(self):
        self._c("snapshot_save")

    def snapshot_restore(self):
        self._c("snapshot_restore")

    def remove_mean_velocity(self):
        self._c("remove_mean_velocity")

    def scale_to_temperature(self, t):
        self._c("scale_to_temperature", float(t))

    def jam(self, params: JamParams | None = None) -> JamResult:
        res = JamResult()
        self._c("jam", C.byref(params or jam_params()), C.byref(res))
        return res

    def bumpy_setup(self, n_vertices_per_particle, vertex_radius, rotation=True):
        a = np.ascontiguousarray(n_vertices_per_particle, dtype=np.int64)
        self._c("bumpy_setup", _ptr(a), float(vertex_radius), int(bool(rotation)))

    # GPU-only execution controls
    def set_option(self, opt, value):
        self._c("set_option", int(opt), int(value))

    def kernel_time(self, cls):
        ms, n = C.c_double(), C.c_int64()
        self._c("kernel_time", int(cls), C.byref(ms), C.byref(n))
        return ms.value, n.value

    def reset_kernel_time(self):
        self._c("reset_kernel_time")

    def synchronize(self):
        self._c("synchronize")


class Batch:
    """Many independent systems jammed together on one GPU (dpm_batch_*): disks, and with
    `vertices` (vertices per system, 0 for a disk system) rigid bumpy systems beside them."""

    def __init__(self, sizes, vertices=None):
        self.be = load()
        self.sizes = [int(n) for n in sizes]
        a = np.ascontiguousarray(self.sizes, dtype=np.int64)
        if vertices is None:
            self.vertices = [0] * len(self.sizes)
            self.h = self.be.fn["batch_create"](len(self.sizes), _ptr(a))
        else:
            self.vertices = [int(v) for v in vertices]
            v = np.ascontiguousarray(self.vertices, dtype=np.int64)
            self.h = self.be.fn["batch_create_mixed"](len(self.sizes), _ptr(a), _ptr(v))
        if not self.h:
            raise DpmError(f"dpm_batch_create: {self.be.err()}")

    def close(self):
        if self.h:
            self.be.fn["batch_destroy"](self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _c(self, name, *args):
        r = self.be.fn[name](self.h, *args)
        if r < 0:
            raise DpmError(f"dpm_{name}: {self.be.err()}")
        return r

    def set_system(self, k, x, y, rad, mass, side):
        if not 0 <= k < len(self.sizes):
            raise DpmError(f"set_system: no system {k}")
        arrs = [np.ascontiguousarray(v, dtype=np.float64) for v in (x, y, rad, mass)]
        if any(len(v) != self.sizes[k] for v in arrs):
            raise DpmError("set_system: array lengths differ from the system size")
        self._c("batch_set_system", int(k), *[_ptr(v) for v in arrs], float(side))

    def get_system(self, k):
        x = np.empty(self.sizes[k])
        y = np.empty(self.sizes[k])
        side = C.c_double()
        self._c("batch_get_system", int(k), _ptr(x), _ptr(y), C.byref(side))
        return x, y, side.value

    def set_rigid_system(self, k, x, y, rad, mass, angle, vertices_per_particle, vertex_radius,
                         side, e_c=1.0, rotation=True):
        if not 0 <= k < len(self.sizes):
            raise DpmError(f"set_rigid_system: no system {k}")
        arrs = [np.ascontiguousarray(v, dtype=np.float64) for v in (x, y, rad, mass, angle)]
        nv = np.ascontiguousarray(vertices_per_particle, dtype=np.int64)
        if any(len(v) != self.sizes[k] for v in arrs) or len(nv) != self.sizes[k]:
            raise DpmError("set_rigid_system: array lengths differ from the system size")
        self._c("batch_set_rigid_system", int(k), *[_ptr(v) for v in arrs], _ptr(nv),
                float(vertex_radius), float(e_c), int(bool(rotation)), float(side))

    def get_rigid_system(self, k):
        """(x, y, angle, vertex_x, vertex_y, box side) of rigid system k."""
        n, nv = self.sizes[k], self.vertices[k]
        x, y, a = np.empty(n), np.empty(n), np.empty(n)
        vx, vy = np.empty(nv), np.empty(nv)
        side = C.c_double()
        self._c("batch_get_rigid_system", int(k), _ptr(x), _ptr(y), _ptr(a), _ptr(vx), _ptr(vy),
                C.byref(side))
        return x, y, a, vx, vy, side.value

    def config(self, e_c=1.0, n_c=2.0, cutoff=1.7, disp_thresh_sq=3 * 0.15 ** 2):
        self._c("batch_config", float(e_c), float(n_c), float(cutoff), float(disp_thresh_sq))

    def set_option(self, opt, value):
        self._c("batch_set_option", int(opt), int(value))

    def autotune(self, params: JamParams, trial_steps=200, cutoffs=(1.6, 1.7, 1.9),
                 chunks=(16, 32, 64)):
        c = np.ascontiguousarray(cutoffs, dtype=np.float64)
        k = np.ascontiguousarray(chunks, dtype=np.int64)
        bc, bk, ms = C.c_double(), C.c_int64(), C.c_double()
        self._c("batch_autotune", C.byref(params), int(trial_steps), _ptr(c), len(c), _ptr(k),
                len(k), C.byref(bc), C.byref(bk), C.byref(ms))
        return bc.value, bk.value, ms.value

    def autotune_shape(self, params: JamParams, trial_steps=200):
        """Force-pass workgroup size and cell width (dpm_batch_autotune_shape)."""
        blk, pct, ms = C.c_int64(), C.c_int64(), C.c_double()
        self._c("batch_autotune_shape", C.byref(params), int(trial_steps), C.byref(blk),
                C.byref(pct), C.byref(ms))
        return blk.value, pct.value, ms.value

    def kernel_time(self):
        """(milliseconds, launches) of the timed force passes (DPM_OPT_TIMING)."""
        ms, n = C.c_double(), C.c_int64()
        self._c("batch_kernel_time", C.byref(ms), C.byref(n))
        return ms.value, n.value

    def row_entries(self):
        t = C.c_int64()
        self._c("batch_row_entries", C.byref(t))
        return t.value

    def vertex_row_entries(self):
        t = C.c_int64()
        self._c("batch_vertex_row_entries", C.byref(t))
        return t.value

    def finish_steps(self):
        out = np.empty(len(self.sizes), dtype=np.int64)
        self._c("batch_finish_steps", _ptr(out))
        return out

    def jam(self, params: JamParams, max_batched_steps=10**9):
        out = (JamResult * len(self.sizes))()
        steps = C.c_int64()
        self._c("batch_jam", C.byref(params), int(max_batched_steps), C.cast(out, C.c_void_p),
                C.byref(steps))
        return list(out), steps.value

    def minimize(self, params: FireParams, max_batched_steps=10**9):
        """minimizeFire of every system at once (dpm_batch_minimize)."""
        out = (JamResult * len(self.sizes))()
        steps = C.c_int64()
        self._c("batch_minimize", C.byref(params), int(max_batched_steps),
                C.cast(out, C.c_void_p), C.byref(steps))
        return list(out), steps.value


REQ_JAM = 0
REQ_MINIMIZE = 1


class Server:
    """Continuous batching of jamming and minimisation requests on one GPU (dpm_server_*)."""

    def __init__(self, max_systems=4096, arena_slots=1 << 22, budget_bytes=8 << 30,
                 vertex_slots=1 << 22):
        self.be = load()
        self.h = self.be.fn["server_create"](int(max_systems), int(arena_slots),
                                             int(budget_bytes), int(vertex_slots))
        if not self.h:
            raise DpmError(f"dpm_server_create: {self.be.err()}")
        self.n = {}

    def close(self):
        if self.h:
            self.be.fn["server_destroy"](self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, x, y, rad, mass, side, jam: JamParams, e_c=1.0, angle=None,
               vertices_per_particle=None, vertex_radius=0.0, mode=REQ_JAM):
        """mode REQ_JAM: dpm_jam with `jam`; REQ_MINIMIZE: minimizeFire with jam.fire."""
        arrs = [np.ascontiguousarray(v, dtype=np.float64) for v in (x, y, rad, mass)]
        r = Request()
        bumpy = angle is not None
        r.kind = BUMPY if bumpy else DISK
        r.mode = int(mode)
        r.n_particles = len(arrs[0])
        r.x, r.y, r.radius, r.mass = [_ptr(a) for a in arrs]
        keep = list(arrs)
        if bumpy:
            a = np.ascontiguousarray(angle, dtype=np.float64)
            v = np.ascontiguousarray(vertices_per_particle, dtype=np.int64)
            keep += [a, v]
            r.angle, r.vertices_per_particle = _ptr(a), _ptr(v)
            r.n_vertices = int(v.sum())
            r.vertex_radius = float(vertex_radius)
        r.box_side = float(side)
        r.e_c = float(e_c)
        r.jam = jam
        rid = self.be.fn["server_submit"](self.h, C.byref(r))
        if rid < 0:
            raise DpmError(f"dpm_server_submit: {self.be.err()}")
        self.n[rid] = r.n_particles
        return rid

    def progress(self, batched_steps=32):
        f, i, q = C.c_int64(), C.c_int64(), C.c_int64()
        if self.be.fn["server_progress"](self.h, int(batched_steps), C.byref(f), C.byref(i),
                                         C.byref(q)) < 0:
            raise DpmError(f"dpm_server_progress: {self.be.err()}")
        return f.value, i.value, q.value

    def result(self, rid):
        st = C.c_int32()
        out = JamResult()
        lat = C.c_double()
        n = self.n.get(rid, 0)  # unknown (or released) ids are reported by the library
        x, y = np.empty(n), np.empty(n)
        side = C.c_double()
        if self.be.fn["server_result"](self.h, int(rid), C.byref(st), C.byref(out), C.byref(lat),
                                       _ptr(x) if n else None, _ptr(y) if n else None,
                                       C.byref(side)) < 0:
            raise DpmError(f"dpm_server_result: {self.be.err()}")
        if st.value != 2:
            return {"state": st.value}
        return {"state": 2, "result": out, "latency_s": lat.value, "x": x, "y": y,
                "box_side": side.value}

    def release(self, rid):
        """Forget a finished request (dpm_server_release)."""
        if self.be.fn["server_release"](self.h, int(rid)) < 0:
            raise DpmError(f"dpm_server_release: {self.be.err()}")
        self.n.pop(rid, None)
