"""GPU test of the allocation-free entry point dpm_disk_forces_noalloc, called through the
C ABI on caller-owned device arrays: its forces, energies and total energy equal those of
the list-based entry points (dpm_neighbor_init with the Verlet method, dpm_zero_forces,
dpm_compute_forces, dpm_reduce(DPM_RED_PE)) on the same disks, bit for bit, on calls in a
row that share the library's scratch array, the later ones with fewer particles."""
import ctypes as C

import numpy as np
import pytest

from _helpers import pkg, workloads

P = pkg()
W = workloads()
pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def say(*a):
    print("[gpu-case]", *a, flush=True)


class Hip:
    """Device buffers for the test's inputs and outputs (the runtime libdpm.so uses)."""

    def __init__(self):
        P.load()
        self.lib = C.CDLL("libamdhip64.so.7")
        self.lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.lib.hipFree.argtypes = [C.c_void_p]
        self.lib.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.lib.hipStreamSynchronize.argtypes = [C.c_void_p]
        self.lib.hipStreamDestroy.argtypes = [C.c_void_p]
        self.held = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.lib.hipMalloc(C.byref(p), max(int(nbytes), 8)) == 0
        self.held.append(p)
        return p

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = self.alloc(a.nbytes)
        assert self.lib.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, H2D) == 0
        return p

    def get(self, p, n):
        out = np.empty(n)
        assert self.lib.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, D2H) == 0
        return out

    def free(self):
        for p in self.held:
            self.lib.hipFree(p)
        self.held = []


def noalloc(hip, stream, st, cutoff, e_c=1.0, n_c=2.0):
    n = st.n
    x, y, r = hip.put(st.x), hip.put(st.y), hip.put(st.rad)
    fx, fy, pe, tot = hip.alloc(8 * n), hip.alloc(8 * n), hip.alloc(8 * n), hip.alloc(8)
    fn = P.load().fn["disk_forces_noalloc"]
    rc = fn(x, y, r, n, st.side, st.side, e_c, n_c, cutoff, fx, fy, pe, tot, stream)
    assert rc == 0, P.load().err()
    assert hip.lib.hipStreamSynchronize(stream) == 0
    return hip.get(fx, n), hip.get(fy, n), hip.get(pe, n), hip.get(tot, 1)[0]


def listed(st, cutoff, e_c=1.0, n_c=2.0):
    spec = W.NeighborSpec(2, cutoff, 3 * 0.15 ** 2, 3, 1e30)
    s = W.make_disk_system(P.System, st, spec, e_c=e_c, n_c=n_c)
    s.zero_forces()
    s.compute_forces()
    out = (s.download(P.F_FX), s.download(P.F_FY), s.download(P.F_PE), s.reduce(P.RED_PE))
    s.close()
    return out


def test_noalloc_forces_equal_the_list_path_on_calls_in_a_row():
    hip = Hip()
    stream = C.c_void_p()
    assert hip.lib.hipStreamCreate(C.byref(stream)) == 0
    cases = [(W.disk_config(6000, 0.86, seed=71), 1.7, 1.0, 2.0),   # cell grid
             (W.disk_config(700, 0.84, seed=72), 1.7, 1.0, 2.0),    # fewer particles
             (W.disk_config(12, 0.6, seed=73), 1.7, 1.0, 2.0),      # box under 3 cutoffs
             (W.disk_config(300, 0.88, seed=74), 2.1, 0.5, 2.0)]    # other cutoff and e_c
    for st, cutoff, e_c, n_c in cases:
        say("noalloc", st.n, st.side, cutoff, e_c)
        got = noalloc(hip, stream, st, cutoff, e_c, n_c)
        ref = listed(st, cutoff, e_c, n_c)
        for k, name in enumerate(("fx", "fy", "pe")):
            assert np.array_equal(got[k], ref[k]), (st.n, name, np.max(np.abs(got[k] - ref[k])))
        assert got[3] == ref[3], (st.n, got[3], ref[3])
        assert np.max(np.abs(got[0])) > 0  # the systems overlap: the comparison is not of zeros
    hip.lib.hipStreamDestroy(stream)
    hip.free()


def test_noalloc_rejects_sizes_beyond_its_scratch():
    hip = Hip()
    fn = P.load().fn["disk_forces_noalloc"]
    buf = hip.alloc(64)
    for n in (0, (1 << 20) + 1):
        rc = fn(buf, buf, buf, n, 10.0, 10.0, 1.0, 2.0, 1.7, buf, buf, buf, buf, None)
        assert rc < 0 and "1048576" in P.load().err()
    st = W.disk_config(5000, 0.9, seed=75)  # box under three cutoffs a side, too many disks
    rc = fn(buf, buf, buf, st.n, 4.0, 4.0, 1.0, 2.0, 1.7, buf, buf, buf, buf, None)
    assert rc < 0 and "4096" in P.load().err()
    hip.free()
