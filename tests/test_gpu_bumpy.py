"""GPU parity for the rigid bumpy path (rings of vertex disks): libdpm.so against the
CPU oracle.

Vertex rings and rigid-body rotations use one shared sin/cos (dpm_sincos) on both sides,
so set-up, vertex and particle pair sets, forces, torques, energies, NVE and FIRE
trajectories, Adam, stress without spin and contacts agree bit for bit.  The stress tensor of
spinning particles goes through the library atan2/sin/cos and is held to 1e-12.
The parity bars (forces 1e-12, NVE 1e-9 over 100 steps, FIRE final E, P, phi 1e-8)
are therefore met with margin.
"""
import numpy as np
import pytest

from _helpers import oracle, pkg, workloads

P = pkg()
W = workloads()
pytestmark = pytest.mark.gpu

PART = ("F_X", "F_Y", "F_FX", "F_FY", "F_PE", "F_TORQUE", "F_ANGLE")


def say(*a):
    print("[gpu-case]", *a, flush=True)


def bpair(st, method=3, rotation=True, same_vertices=False):
    g = W.make_bumpy_system(P.System, st, method=method, rotation=rotation)
    o = W.make_bumpy_system(P.System, st, backend=oracle(), method=method, rotation=rotation)
    if same_vertices:
        g.upload(P.F_VERTEX_X, o.download(P.F_VERTEX_X))
        g.upload(P.F_VERTEX_Y, o.download(P.F_VERTEX_Y))
        g.neighbor_init()
        o.neighbor_init()
    return g, o


def forces(s):
    s.zero_forces()
    s.compute_forces()


def same(g, o, fields, rtol=0.0, atol=0.0):
    for f in fields:
        a, b = g.download(getattr(P, f)), o.download(getattr(P, f))
        if rtol == 0.0 and atol == 0.0:
            assert np.array_equal(a, b), (f, np.max(np.abs(a - b)))
        else:
            np.testing.assert_allclose(a, b, rtol=rtol, atol=atol, err_msg=f)


def sorted_pairs(i, j):
    k = np.lexsort((j, i))
    return i[k], j[k]


def test_geometry_masses_and_inertia():
    st = W.bumpy_config(300, 0.7, seed=4)
    g, o = bpair(st)
    for f in ("F_VERTEX_X", "F_VERTEX_Y", "F_VERTEX_MASS"):
        assert np.array_equal(g.download(getattr(P, f)), o.download(getattr(P, f))), f
    assert np.array_equal(g.download(P.F_NUM_VERTICES), st.nverts)
    assert np.array_equal(g.download(P.F_MOI), o.download(P.F_MOI))


@pytest.mark.parametrize("n,phi,method", [(64, 0.8, 2), (512, 0.75, 3), (2000, 0.82, 3),
                                          (700, 0.6, 1)])
def test_vertex_pairs_and_forces_bit_exact(n, phi, method):
    say("bumpy pairs/forces", n, phi, method)
    st = W.bumpy_config(n, phi, seed=n)
    g, o = bpair(st, method=method)
    gi, gj = sorted_pairs(*g.neighbor_pairs())
    oi, oj = sorted_pairs(*o.neighbor_pairs())
    assert np.array_equal(gi, oi) and np.array_equal(gj, oj)
    gi, gj = sorted_pairs(*g.vertex_pairs())
    oi, oj = sorted_pairs(*o.vertex_pairs())
    assert len(gi) > 0
    assert np.array_equal(gi, oi) and np.array_equal(gj, oj)
    forces(g)
    forces(o)
    same(g, o, PART)
    assert g.reduce(P.RED_PE) == o.reduce(P.RED_PE)
    # Newton's third law over vertex pairs: the net force vanishes
    assert abs(g.reduce(P.RED_FORCE_BALANCE)) / n < 1e-14


def test_stress_pressure_and_contacts():
    st = W.bumpy_config(800, 0.85, seed=8)
    g, o = bpair(st, same_vertices=True)
    for s in (g, o):
        forces(s)
        s.compute_stress()
        s.count_contacts()
    same(g, o, ("F_STRESS_XX", "F_STRESS_XY", "F_STRESS_YX", "F_STRESS_YY"))
    assert np.array_equal(g.download(P.F_CONTACTS), o.download(P.F_CONTACTS))
    assert np.array_equal(g.download(P.F_VERTEX_CONTACTS), o.download(P.F_VERTEX_CONTACTS))
    assert g.reduce(P.RED_PRESSURE) == o.reduce(P.RED_PRESSURE)
    assert g.reduce(P.RED_CONTACTS) == o.reduce(P.RED_CONTACTS) > 0


def test_stress_of_spinning_particles():
    # the rotational velocity term goes through the library atan2/sin/cos on each side
    st = W.bumpy_config(400, 0.8, seed=10)
    g, o = bpair(st)
    w = np.random.default_rng(4).normal(0, 0.3, st.n)
    for s in (g, o):
        s.upload(P.F_OMEGA, w)
        forces(s)
        s.compute_stress()
    same(g, o, ("F_STRESS_XX", "F_STRESS_XY", "F_STRESS_YX", "F_STRESS_YY"), rtol=1e-12,
         atol=1e-15)
    assert g.reduce(P.RED_PRESSURE) == pytest.approx(o.reduce(P.RED_PRESSURE), rel=1e-12)


@pytest.mark.parametrize("rotation", [True, False])
def test_nve_trajectory(rotation):
    st = W.bumpy_config(500, 0.75, seed=21)
    g, o = bpair(st, rotation=rotation)
    rng = np.random.default_rng(3)
    vx, vy = rng.normal(0, 1e-2, st.n), rng.normal(0, 1e-2, st.n)
    w = rng.normal(0, 1e-2, st.n) if rotation else np.zeros(st.n)
    for s in (g, o):
        s.upload(P.F_VX, vx)
        s.upload(P.F_VY, vy)
        s.upload(P.F_OMEGA, w)
        forces(s)
    e0 = g.reduce(P.RED_PE) + g.reduce(P.RED_KE)
    g.nve_run(0.002, 100)
    o.nve_run(0.002, 100)
    same(g, o, ("F_X", "F_Y", "F_VX", "F_VY", "F_ANGLE", "F_OMEGA"))
    same(g, o, ("F_FX", "F_FY", "F_TORQUE", "F_PE"))
    if not rotation:
        assert np.array_equal(g.download(P.F_OMEGA), np.zeros(st.n))
        assert np.array_equal(g.download(P.F_ANGLE), st.angle)
    e1 = g.reduce(P.RED_PE) + g.reduce(P.RED_KE)
    assert abs(e1 / e0 - 1) < 1e-3


def test_cell_reorder_during_nve():
    # fast particles cross cells: the cell sort permutes particles and their vertex blocks
    st = W.bumpy_config(400, 0.6, seed=5)
    g, o = bpair(st, method=3)
    rng = np.random.default_rng(11)
    vx, vy = rng.normal(0, 0.5, st.n), rng.normal(0, 0.5, st.n)
    for s in (g, o):
        s.upload(P.F_VX, vx)
        s.upload(P.F_VY, vy)
        forces(s)
    g.nve_run(0.01, 150)
    o.nve_run(0.01, 150)
    assert g.neighbor_stats()["n_cell_rebuilds"] == o.neighbor_stats()["n_cell_rebuilds"] > 0
    same(g, o, ("F_X", "F_Y", "F_VX", "F_VY", "F_ANGLE", "F_FX", "F_FY", "F_TORQUE"))
    for f in ("F_VERTEX_X", "F_VERTEX_Y"):
        assert np.array_equal(g.download(getattr(P, f)), o.download(getattr(P, f))), f


def test_fire_to_max_force():
    # random rings overlap deeply; at phi = 0.6 FIRE reaches |F|max < 1e-12 in ~1000 steps
    st = W.bumpy_config(128, 0.6, seed=2)
    g, o = bpair(st)
    prm = P.fire_params(avg_pe_target=0.0, avg_pe_diff_target=0.0, fmax_target=1e-12,
                        max_steps=200000)
    rg, ro = g.minimize_fire(prm), o.minimize_fire(prm)
    say("bumpy fire", rg.steps, ro.steps, rg.converged, ro.converged)
    assert rg.converged == ro.converged == 1 and rg.steps == ro.steps
    for s in (g, o):
        s.compute_stress()
    for what in (P.RED_PE, P.RED_PRESSURE, P.RED_PACKING_FRACTION, P.RED_FMAX_SQ):
        assert g.reduce(what) == o.reduce(what), what
    same(g, o, PART)


def test_fire_step_trace_with_rotation():
    st = W.bumpy_config(256, 0.85, seed=17)
    g, o = bpair(st)
    for s in (g, o):
        s.zero_velocities()
        forces(s)
    sg, so = g.fire_state(0.1, 0.01), o.fire_state(0.1, 0.01)
    for k in range(40):
        g.fire_step(sg)
        o.fire_step(so)
        assert (sg.power, sg.dt, sg.alpha, sg.n_good, sg.n_bad, sg.stopped) == \
            (so.power, so.dt, so.alpha, so.n_good, so.n_bad, so.stopped), k
    same(g, o, PART)


@pytest.mark.parametrize("rotation", [True, False])
def test_adam_matches_oracle(rotation):
    # rigid Adam (kernelRigidBumpyAdamStep + the vertex pass) through minimizeAdam's loop;
    # 400 steps of up to 2e-3 per coordinate move particles across cells, so the angle
    # moments are permuted with their particles at the cell reorders
    say("bumpy adam", rotation)
    st = W.bumpy_config(256, 0.85, seed=23)
    g, o = bpair(st, rotation=rotation)
    ng = g.minimize_adam(alpha=2e-3, max_steps=400)
    no = o.minimize_adam(alpha=2e-3, max_steps=400)
    assert ng == no
    assert g.neighbor_stats()["n_cell_rebuilds"] == o.neighbor_stats()["n_cell_rebuilds"] > 0
    same(g, o, PART)
    for f in ("F_VERTEX_X", "F_VERTEX_Y"):
        assert np.array_equal(g.download(getattr(P, f)), o.download(getattr(P, f))), f


def test_snapshot_restore():
    st = W.bumpy_config(300, 0.8, seed=9)
    g, o = bpair(st)
    for s in (g, o):
        forces(s)
        s.snapshot_save()
        s.scale_box(0.98)
        s.minimize_fire(P.fire_params(max_steps=30))
        s.snapshot_restore()
        forces(s)
    same(g, o, PART)
    for f in ("F_VERTEX_X", "F_VERTEX_Y"):
        assert np.array_equal(g.download(getattr(P, f)), o.download(getattr(P, f))), f


def test_n65536_rigid_size_properties():
    n = 65536
    say("rigid benchmark size", n)
    st = W.bumpy_config(n, 0.8, seed=12345)
    g, o = bpair(st)
    gi, gj = sorted_pairs(*g.vertex_pairs())
    oi, oj = sorted_pairs(*o.vertex_pairs())
    assert np.array_equal(gi, oi) and np.array_equal(gj, oj)
    forces(g)
    forces(o)
    same(g, o, ("F_FX", "F_FY", "F_PE", "F_TORQUE"))
    assert abs(g.reduce(P.RED_FORCE_BALANCE)) / n < 1e-14
