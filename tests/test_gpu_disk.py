"""GPU parity for the disk hot path: libdpm.so (gfx950 HIP) against the CPU oracle.

Bar: neighbour pair sets identical; forces, energies, NVE trajectories, FIRE and
jamming results bit-identical where the arithmetic is the same IEEE sequence on both
sides (disk path: no transcendental functions, no FMA contraction, same reduction
tree), and at minimum within the stated tolerances (forces/energies 1e-12 relative,
NVE 1e-9 over 100 steps, FIRE/jamming final E, P, phi 1e-8).
"""
import numpy as np
import pytest

from _helpers import ROOT, oracle, pkg, workloads

P = pkg()
W = workloads()
pytestmark = pytest.mark.gpu

FIELDS = ("F_X", "F_Y", "F_VX", "F_VY", "F_FX", "F_FY", "F_PE")


def say(*a):
    print("[gpu-case]", *a, flush=True)


def pair(st, spec=None, method=None):
    spec = spec or W.neighbor_spec(st.side, method=method or 3)
    g = W.make_disk_system(P.System, st, spec)
    o = W.make_disk_system(P.System, st, spec, backend=oracle())
    return g, o


def assert_same_state(g, o, fields=FIELDS, rtol=0.0, atol=0.0):
    for f in fields:
        a, b = g.download(getattr(P, f)), o.download(getattr(P, f))
        if rtol == 0.0 and atol == 0.0:
            assert np.array_equal(a, b), (f, np.max(np.abs(a - b)))
        else:
            np.testing.assert_allclose(a, b, rtol=rtol, atol=atol, err_msg=f)


def forces(s):
    s.zero_forces()
    s.compute_forces()


# ------------------------------------------------------------------ fixture
def test_golden_fixture_n1024():
    z = np.load(f"{ROOT}/tests/golden/disk_n1024.npz")
    say("golden N=1024 phi=0.80")
    st = W.DiskState(1024, float(z["side"]), z["x"], z["y"], z["rad"], z["mass"])
    spec = W.NeighborSpec(3, float(z["cutoff"]), float(z["disp_thresh_sq"]),
                          int(z["n_cells_dim"]), float(z["cell_disp_thresh_sq"]))
    g = W.make_disk_system(P.System, st, spec)
    g.upload(P.F_VX, z["vx"])
    g.upload(P.F_VY, z["vy"])
    forces(g)
    assert np.array_equal(g.download(P.F_STATIC_INDEX), z["static_index"])
    gi, gj = g.neighbor_pairs()
    assert np.array_equal(gi, z["pair_i"]) and np.array_equal(gj, z["pair_j"])
    np.testing.assert_allclose(g.download(P.F_FX), z["fx"], rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(g.download(P.F_FY), z["fy"], rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(g.download(P.F_PE), z["pe"], rtol=1e-12, atol=1e-300)
    assert g.reduce(P.RED_PRESSURE) == pytest.approx(float(z["pressure"]), rel=1e-12)
    assert g.reduce(P.RED_CONTACTS) == float(z["contacts"])
    g.nve_run(0.01, 100)
    for f, k in ((P.F_X, "x100"), (P.F_Y, "y100"), (P.F_VX, "vx100"), (P.F_VY, "vy100")):
        np.testing.assert_allclose(g.download(f), z[k], rtol=1e-9, atol=1e-12)
    assert g.reduce(P.RED_PE) == pytest.approx(float(z["pe100"]), rel=1e-9)


# ------------------------------------------------------------------ neighbours + forces
@pytest.mark.parametrize("n,phi,method", [(1024, 0.80, 3), (4096, 0.85, 3), (600, 0.8, 2),
                                          (200, 0.9, 1), (3000, 0.3, 3), (2500, 0.82, 2),
                                          (1024, 0.3, 2)])
def test_pairs_and_forces_bit_exact(n, phi, method):
    say("pairs+forces", n, phi, method)
    st = W.disk_config(n, phi, seed=n + 7)
    if method == 1:
        g = W.make_disk_system(P.System, st, W.NeighborSpec(1, 0, 0, 0, 0))
        o = W.make_disk_system(P.System, st, W.NeighborSpec(1, 0, 0, 0, 0), backend=oracle())
    else:
        g, o = pair(st, method=method)
    assert np.array_equal(g.download(P.F_STATIC_INDEX), o.download(P.F_STATIC_INDEX))
    gi, gj = g.neighbor_pairs()
    oi, oj = o.neighbor_pairs()
    assert np.array_equal(gi, oi) and np.array_equal(gj, oj)
    assert g.neighbor_stats()["max_alloc"] == o.neighbor_stats()["max_alloc"]
    forces(g)
    forces(o)
    assert_same_state(g, o)
    for red in (P.RED_PE, P.RED_FORCE_BALANCE, P.RED_PRESSURE, P.RED_CONTACTS, P.RED_KE,
                P.RED_PACKING_FRACTION, P.RED_FMAX_SQ):
        assert g.reduce(red) == o.reduce(red), red


def test_general_exponent_forces_and_nve():
    # n_c != 2 takes the general pair path, whose overlap**(n_c - 1) is the device math
    # library's pow(); it and the host's pow() are not correctly rounded and can differ
    # in the last bit, so this case is held to the stated tolerances rather than bits.
    st = W.disk_config(1500, 0.86, seed=17)
    g, o = pair(st)
    vx, vy = W.gaussian_velocities(st.n, 1e-3, 8)
    for s in (g, o):
        s.set_energy(0.7, 2.5)
        s.upload(P.F_VX, vx)
        s.upload(P.F_VY, vy)
        forces(s)
    assert_same_state(g, o, fields=("F_FX", "F_FY", "F_PE"), rtol=1e-12, atol=1e-15)
    for s in (g, o):
        s.nve_run(0.01, 150)
    assert_same_state(g, o, fields=("F_X", "F_Y", "F_VX", "F_VY"), rtol=1e-9, atol=1e-12)
    assert g.reduce(P.RED_PRESSURE) == pytest.approx(o.reduce(P.RED_PRESSURE), rel=1e-9)


def test_stress_tensor_components():
    st = W.disk_config(2000, 0.86, seed=3)
    g, o = pair(st)
    vx, vy = W.gaussian_velocities(st.n, 1e-3, 5)
    for s in (g, o):
        s.upload(P.F_VX, vx)
        s.upload(P.F_VY, vy)
        s.compute_stress()
    for f in ("F_STRESS_XX", "F_STRESS_XY", "F_STRESS_YX", "F_STRESS_YY"):
        assert np.array_equal(g.download(getattr(P, f)), o.download(getattr(P, f))), f


# ------------------------------------------------------------------ NVE
@pytest.mark.parametrize("n,phi,T,steps", [(1024, 0.80, 1e-4, 100), (2048, 0.84, 5e-2, 300),
                                           (512, 0.7, 1.0, 200)])
def test_nve_trajectory_matches_oracle(n, phi, T, steps):
    say("nve", n, phi, T, steps)
    st = W.disk_config(n, phi, seed=n)
    g, o = pair(st)
    vx, vy = W.gaussian_velocities(n, T, n)
    for s in (g, o):
        s.upload(P.F_VX, vx)
        s.upload(P.F_VY, vy)
        forces(s)
        s.nve_run(0.01, steps)
    sg, so = g.neighbor_stats(), o.neighbor_stats()
    assert sg["n_rebuilds"] == so["n_rebuilds"]
    assert sg["n_cell_rebuilds"] == so["n_cell_rebuilds"]
    assert_same_state(g, o)
    e_g = g.reduce(P.RED_PE) + g.reduce(P.RED_KE)
    e_o = o.reduce(P.RED_PE) + o.reduce(P.RED_KE)
    assert e_g == e_o


@pytest.mark.parametrize("steps", [333, 334])
def test_nve_chunking_and_graphs_do_not_change_results(steps):
    # 334 steps leave an odd remainder for every even chunk length (split last step)
    st = W.disk_config(3000, 0.8, seed=44)
    spec = W.neighbor_spec(st.side)
    vx, vy = W.gaussian_velocities(st.n, 0.3, 4)
    o = W.make_disk_system(P.System, st, spec, backend=oracle())
    o.upload(P.F_VX, vx)
    o.upload(P.F_VY, vy)
    forces(o)
    o.nve_run(0.01, steps)
    so = o.neighbor_stats()
    assert so["n_rebuilds"] > 1 and so["n_cell_rebuilds"] > 1  # the run crossed rebuilds
    for graph, chunk in ((1, 64), (0, 64), (1, 2), (0, 10), (1, 256)):
        say("chunking", steps, graph, chunk)
        g = W.make_disk_system(P.System, st, spec)
        g.set_option(P.OPT_GRAPH, graph)
        g.set_option(P.OPT_CHUNK, chunk)
        g.upload(P.F_VX, vx)
        g.upload(P.F_VY, vy)
        forces(g)
        g.nve_run(0.01, steps)
        assert_same_state(g, o)
        sg = g.neighbor_stats()
        assert (sg["n_rebuilds"], sg["n_cell_rebuilds"]) == (so["n_rebuilds"],
                                                             so["n_cell_rebuilds"])


def test_short_runs_and_closing_halts_match_oracle():
    """Runs of every short length back to back, at a temperature where the list and the
    cells are rebuilt every few steps, so that the last step of many runs (the closing
    force, a unit of its own) needs a rebuild itself; graphs on and off."""
    st = W.disk_config(1500, 0.8, seed=45)
    spec = W.neighbor_spec(st.side)
    vx, vy = W.gaussian_velocities(st.n, 0.5, 5)
    lengths = [1, 1, 2, 3, 1, 5, 19, 20, 33, 1, 1, 7, 2, 31, 32, 1, 64, 17, 1, 1]
    o = W.make_disk_system(P.System, st, spec, backend=oracle())
    o.upload(P.F_VX, vx)
    o.upload(P.F_VY, vy)
    forces(o)
    for n in lengths:
        o.nve_run(0.01, n)
    so = o.neighbor_stats()
    assert so["n_cell_rebuilds"] > 5
    for graph, chunk in ((1, 32), (0, 32), (1, 8)):
        say("short runs", graph, chunk)
        g = W.make_disk_system(P.System, st, spec)
        g.set_option(P.OPT_GRAPH, graph)
        g.set_option(P.OPT_CHUNK, chunk)
        g.upload(P.F_VX, vx)
        g.upload(P.F_VY, vy)
        forces(g)
        for n in lengths:
            g.nve_run(0.01, n)
        assert_same_state(g, o)
        sg = g.neighbor_stats()
        assert (sg["n_rebuilds"], sg["n_cell_rebuilds"]) == (so["n_rebuilds"],
                                                             so["n_cell_rebuilds"])


def test_verlet_grid_rows_through_nve_and_jam():
    """The Verlet method builds the rows of systems of 1,024+ particles from a cell grid
    (the nine cells' lists merged in ascending index): the rows, and so every bit of an
    NVE run with rebuilds and of a jam whose box changes at every compression step, equal
    the oracle's all-pairs Verlet rows."""
    st = W.disk_config(2000, 0.80, seed=46)
    g, o = pair(st, method=2)
    vx, vy = W.gaussian_velocities(st.n, 0.05, 6)
    for s in (g, o):
        s.upload(P.F_VX, vx)
        s.upload(P.F_VY, vy)
        forces(s)
        s.nve_run(0.01, 200)
    sg, so = g.neighbor_stats(), o.neighbor_stats()
    assert sg["n_rebuilds"] == so["n_rebuilds"] > 2
    assert_same_state(g, o)
    gi, gj = g.neighbor_pairs()
    oi, oj = o.neighbor_pairs()
    assert np.array_equal(gi, oi) and np.array_equal(gj, oj)
    prm = P.jam_params(pressure_target=1e-3, window_rel=0.5, strain_step=1e-2,
                       max_compression_steps=30,
                       fire=P.fire_params(avg_pe_target=0.0, avg_pe_diff_target=0.0,
                                          fmax_target=1e-8, max_steps=300))
    for s in (g, o):
        s.upload(P.F_VX, np.zeros(st.n))
        s.upload(P.F_VY, np.zeros(st.n))
    rg, ro = g.jam(prm), o.jam(prm)
    say("verlet grid jam", rg.status, rg.compression_steps, rg.fire_steps_total)
    for f in ("status", "compression_steps", "fire_steps_total", "pressure",
              "packing_fraction", "pe", "box_length", "contacts"):
        assert getattr(rg, f) == getattr(ro, f), f
    assert_same_state(g, o, fields=("F_X", "F_Y"))


def _converging_ring(n_ring=20, speed=1.0, radius=0.28):
    """A centre disk and a ring of disks closing in on it: every ring disk enters the
    centre's cutoff at the same step, so the centre's row jumps from 0 to n_ring
    (beyond the initial 16 slots) inside one list rebuild, while neighbouring ring
    disks press on each other.  Three cells per side make the stencil cover the
    whole box, and the cell threshold is never reached."""
    side = 9.0
    ang = 2 * np.pi * np.arange(n_ring) / n_ring
    x = np.concatenate([[4.5], 4.5 + 2.5 * np.cos(ang)])
    y = np.concatenate([[4.5], 4.5 + 2.5 * np.sin(ang)])
    n = n_ring + 1
    st = W.DiskState(n, side, x, y, np.full(n, radius), np.ones(n))
    spec = W.NeighborSpec(3, 1.7, 3 * 0.15 ** 2, 3, 1e30)
    vx = np.concatenate([[0.0], -speed * np.cos(ang)])
    vy = np.concatenate([[0.0], -speed * np.sin(ang)])
    return st, spec, vx, vy


def test_list_rebuild_overflow_inside_nve():
    st, spec, vx, vy = _converging_ring()
    res = []
    for backend in (None, oracle()):
        s = W.make_disk_system(P.System, st, spec, backend=backend)
        s.upload(P.F_VX, vx)
        s.upload(P.F_VY, vy)
        forces(s)
        assert s.neighbor_stats()["max_alloc"] == 4
        s.nve_run(0.01, 120)
        res.append(s)
    g, o = res
    sg, so = g.neighbor_stats(), o.neighbor_stats()
    say("ring overflow", sg, so)
    assert so["max_alloc"] == 32 and sg["max_alloc"] == 32
    assert sg["n_rebuilds"] == so["n_rebuilds"] and sg["n_cell_rebuilds"] == so["n_cell_rebuilds"]
    assert np.array_equal(g.download(P.F_NUM_NEIGHBORS), o.download(P.F_NUM_NEIGHBORS))
    assert_same_state(g, o)
    assert np.max(np.abs(g.download(P.F_FX))) > 0.01  # the ring is in contact
    gi, gj = g.neighbor_pairs()
    oi, oj = o.neighbor_pairs()
    assert np.array_equal(gi, oi) and np.array_equal(gj, oj)


# ------------------------------------------------------------------ FIRE / Adam
def test_minimize_fire_matches_oracle():
    st = W.disk_config(2048, 0.84, seed=77)
    g, o = pair(st)
    prm = P.fire_params(max_steps=3000)
    rg, ro = g.minimize_fire(prm), o.minimize_fire(prm)
    say("fire steps", rg.steps, ro.steps)
    assert (rg.steps, rg.stopped, rg.converged) == (ro.steps, ro.stopped, ro.converged)
    assert rg.avg_pe == ro.avg_pe and rg.dt == ro.dt and rg.alpha == ro.alpha
    assert_same_state(g, o)


def test_fire_step_trace_matches_oracle():
    st = W.disk_config(700, 0.86, seed=5)
    g, o = pair(st)
    for s in (g, o):
        s.zero_velocities()
        forces(s)
    sg, so = g.fire_state(0.1, 0.01), o.fire_state(0.1, 0.01)
    for k in range(60):
        g.fire_step(sg)
        o.fire_step(so)
        assert (sg.power, sg.dt, sg.alpha, sg.n_good, sg.n_bad, sg.stopped) == \
            (so.power, so.dt, so.alpha, so.n_good, so.n_bad, so.stopped), k
    assert_same_state(g, o)


def test_fire_max_force_stop():
    st = W.disk_config(512, 0.86, seed=9)
    g, o = pair(st)
    prm = P.fire_params(avg_pe_target=0.0, avg_pe_diff_target=0.0, fmax_target=1e-12,
                        max_steps=400000)
    rg, ro = g.minimize_fire(prm), o.minimize_fire(prm)
    assert rg.converged == 1 and rg.fmax_sq < 1e-24
    assert rg.steps == ro.steps
    assert g.reduce(P.RED_PE) == pytest.approx(o.reduce(P.RED_PE), rel=1e-8, abs=1e-30)


def test_minimize_adam_matches_oracle():
    st = W.disk_config(600, 0.8, seed=31)
    g, o = pair(st)
    ng = g.minimize_adam(alpha=1e-3, max_steps=300)
    no = o.minimize_adam(alpha=1e-3, max_steps=300)
    assert ng == no
    assert_same_state(g, o, fields=("F_X", "F_Y", "F_FX", "F_FY", "F_PE"))


# ------------------------------------------------------------------ jamming
def test_jam_small_system_matches_oracle():
    st = W.disk_config(64, 0.80, seed=2)
    g, o = pair(st)
    prm = P.jam_params(pressure_target=1e-4, window_rel=0.05, strain_step=1e-3,
                       max_compression_steps=4000,
                       fire=P.fire_params(avg_pe_target=0.0, avg_pe_diff_target=0.0,
                                          fmax_target=1e-12, max_steps=200000))
    rg, ro = g.jam(prm), o.jam(prm)
    say("jam", rg.status, rg.compression_steps, ro.compression_steps)
    assert rg.status == ro.status == 0
    assert rg.compression_steps == ro.compression_steps
    for k in ("pressure", "packing_fraction", "pe", "box_length"):
        assert getattr(rg, k) == pytest.approx(getattr(ro, k), rel=1e-8), k
    assert 1e-4 <= rg.pressure <= 1.05e-4


# ------------------------------------------------------------------ edge cases
def test_tiny_systems_and_empty_rows():
    for n in (1, 2, 3):
        st = W.disk_config(n, 0.1, seed=n)
        st.side = max(st.side, 6.0)
        g, o = pair(st, method=2)
        forces(g)
        forces(o)
        assert_same_state(g, o)
        g.nve_run(0.01, 10)
        o.nve_run(0.01, 10)
        assert_same_state(g, o)


def test_capacity_regrowth_path():
    # random positions at phi=1.0 give rows longer than the initial capacity
    st = W.disk_config(3000, 1.0, seed=1)
    g, o = pair(st)
    assert g.neighbor_stats()["max_alloc"] == o.neighbor_stats()["max_alloc"] >= 16
    assert np.array_equal(g.download(P.F_NUM_NEIGHBORS), o.download(P.F_NUM_NEIGHBORS))


def test_snapshot_restore_and_scaling():
    st = W.disk_config(1500, 0.8, seed=6)
    g, o = pair(st)
    for s in (g, o):
        forces(s)
        s.snapshot_save()
        s.scale_box(0.97)
        s.minimize_fire(P.fire_params(max_steps=50))
        s.snapshot_restore()
        forces(s)
    assert_same_state(g, o)
    assert g.box() == o.box()


# ------------------------------------------------------------------ full size
def test_n262144_size_properties():
    n = 262144
    say("benchmark size", n)
    st = W.disk_config(n, 0.85, seed=12345)
    g, o = pair(st)
    gi, gj = g.neighbor_pairs()
    oi, oj = o.neighbor_pairs()
    assert np.array_equal(gi, oi) and np.array_equal(gj, oj)
    forces(g)
    forces(o)
    assert_same_state(g, o, fields=("F_FX", "F_FY", "F_PE"))
    assert g.reduce(P.RED_FORCE_BALANCE) / n < 1e-14
    vx, vy = W.gaussian_velocities(n, 1e-5, 3)
    g.upload(P.F_VX, vx)
    g.upload(P.F_VY, vy)
    forces(g)
    g.nve_run(0.001, 200)  # overlapping random state: small dt keeps energy bounded
    e1 = g.reduce(P.RED_PE) + g.reduce(P.RED_KE)
    g.nve_run(0.001, 200)
    e2 = g.reduce(P.RED_PE) + g.reduce(P.RED_KE)
    assert abs(e2 - e1) / abs(e1) < 1e-4
