"""GPU parity for the per-pair analytics (dpm_pair_analytics / dpm_pair_download,
dpm_bumpy_area) against the CPU oracle, through the C ABI.

Every field is compared slot for slot in the reference's layout (rows in the internal cell
order, max_neighbors_allocated slots per row): bit-identical except the separation angle
and the particle area, which use the device library's atan2 / acos and are held to 2 ulp.
Cases: disks with cell and Verlet lists after NVE steps that reorder the particles, a
non-harmonic exponent, the 262,144-disk benchmark size (φ = 0.85, checked against size-independent
properties), and rigid bumpy rings with rotation."""
import numpy as np
import pytest

from _helpers import oracle, pkg, workloads

P = pkg()
W = workloads()
pytestmark = pytest.mark.gpu

ULP_FIELDS = {"pair_separation_angle"}


def say(*a):
    print("[gpu-case]", *a, flush=True)


def compare(g, o):
    a, b = g.pair_analytics(), o.pair_analytics()
    assert set(a) == set(b)
    for k in b:
        assert a[k].shape == b[k].shape, k
        if k in ULP_FIELDS:
            assert np.all(np.abs(a[k] - b[k]) <= 2 * np.spacing(np.abs(b[k]))), k
        else:
            assert np.array_equal(a[k], b[k]), (k, np.argwhere(a[k] != b[k])[:4])
    return a


def run_nve(s, steps, dt=0.01):
    s.zero_forces()
    s.compute_forces()
    s.nve_run(dt, steps)


@pytest.mark.parametrize("n,phi,method,steps", [(500, 0.84, 3, 200), (300, 0.8, 2, 100),
                                                (64, 0.9, 1, 20)])
def test_disk_pairs_match_oracle(n, phi, method, steps):
    say("disk pairs", n, phi, method, steps)
    st = W.disk_config(n, phi, seed=n + 1)
    spec = W.neighbor_spec(st.side, method=method)
    g = W.make_disk_system(P.System, st, spec)
    o = W.make_disk_system(P.System, st, spec, backend=oracle())
    v = W.gaussian_velocities(n, 1e-3, seed=n)
    for s in (g, o):
        s.upload(P.F_VX, v[0])
        s.upload(P.F_VY, v[1])
        run_nve(s, steps)
    assert g.neighbor_stats()["max_alloc"] == o.neighbor_stats()["max_alloc"]
    a = compare(g, o)
    used = a["pair_ids_i"] >= 0
    assert used.sum() == g.neighbor_stats()["total_entries"]


def test_disk_pairs_general_exponent():
    st = W.disk_config(400, 0.86, seed=9)
    g = W.make_disk_system(P.System, st, n_c=2.5)
    o = W.make_disk_system(P.System, st, backend=oracle(), n_c=2.5)
    a, b = g.pair_analytics(), o.pair_analytics()
    # n_c != 2 goes through pow(): the device's and the host's may differ in the last ulp
    for k in b:
        if a[k].dtype == np.int64:
            assert np.array_equal(a[k], b[k]), k
        else:
            np.testing.assert_allclose(a[k], b[k], rtol=1e-13, atol=1e-300, err_msg=k)


def test_disk_pairs_n262144_size_properties():
    """262,144 disks: full symmetric pair set, antisymmetric forces, H_ij = -H_ii exactly,
    fill values in unused slots, and per-particle sums of the slot forces equal to the
    force kernel's result (same row order, so bit-identical)."""
    n = 262144
    st = W.disk_config(n, 0.85, seed=3)
    g = W.make_disk_system(P.System, st)
    g.zero_forces()
    g.compute_forces()
    a = g.pair_analytics()
    ii, jj = a["pair_ids_i"], a["pair_ids_j"]
    used = ii >= 0
    assert used.sum() == g.neighbor_stats()["total_entries"]
    assert np.all(a["potential_pairs"][~used] == -1.0) and np.all(a["force_pairs_x"][~used] == 0)
    key = ii[used] * n + jj[used]
    rkey = jj[used] * n + ii[used]
    order, rorder = np.argsort(key), np.argsort(rkey)
    assert np.array_equal(key[order], rkey[rorder])
    fx = a["force_pairs_x"][used]
    assert np.array_equal(fx[order], -fx[rorder])
    for c in ("xx", "xy", "yx", "yy"):
        assert np.array_equal(a["hessian_pairs_" + c], -a["hessian_ii_" + c]), c
    # slot forces summed along each row in slot order = the particle force
    fsum = np.zeros(n)
    rows = np.where(used, a["force_pairs_x"], 0.0)
    acc = np.zeros(n)
    for k in range(rows.shape[1]):
        acc = acc + rows[:, k]
    sidx = ii[:, 0]
    have = sidx >= 0
    fsum[sidx[have]] = acc[have]
    fx_particles = g.download(P.F_FX)
    assert np.array_equal(fsum[sidx[have]], fx_particles[sidx[have]])


@pytest.mark.parametrize("n,phi,rotation", [(200, 0.75, True), (600, 0.8, False)])
def test_bumpy_pairs_and_area_match_oracle(n, phi, rotation):
    say("bumpy pairs", n, phi, rotation)
    st = W.bumpy_config(n, phi, seed=n + 7)
    g = W.make_bumpy_system(P.System, st, method=3, rotation=rotation)
    o = W.make_bumpy_system(P.System, st, backend=oracle(), method=3, rotation=rotation)
    for s in (g, o):
        s.zero_forces()
        s.compute_forces()
        s.nve_run(0.002, 50)
    compare(g, o)
    ga, oa = g.bumpy_area(), o.bumpy_area()
    assert np.all(np.abs(ga - oa) <= 2 * np.spacing(oa))


def test_pair_download_errors():
    st = W.disk_config(64, 0.8, seed=1)
    g = W.make_disk_system(P.System, st)
    with pytest.raises(P.DpmError):
        g._c("pair_download", 0, 0, 1)  # before dpm_pair_analytics
    g.pair_analytics()
    buf = np.empty(8)
    with pytest.raises(P.DpmError):
        g._c("pair_download", P.PAIR_DISK_FIELDS, buf.ctypes.data, 8)  # bumpy-only field
    with pytest.raises(P.DpmError):
        g._c("pair_download", 0, buf.ctypes.data, 8)  # wrong count
    with pytest.raises(P.DpmError):
        g._c("bumpy_area", buf.ctypes.data, 64)  # not a bumpy system
