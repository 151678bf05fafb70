"""Pins the CPU oracle to what the mathematics fixes.

The reference ships no tests, golden vectors or fixtures, so every oracle function
is checked here against closed forms, invariants or brute force, each chosen so a
plausible slip (a dropped term, a wrong sign or index, a transposed operand) fails:

  pair law (kernels.cuh:154-183)     closed form at one overlap, zero at contact,
                                     F = -dU/dr by finite differences
  pbcDistance (kernels.cuh:65-72)    round-half-away image choice at d = L/2
  cutoff (kernels.cuh:249-260)       strict '<' at exactly rc
  neighbour lists (base/kernels.cu)  cell == verlet == independent brute force
  force sums (disk/kernels.cu:11-57) Newton's third law; PE = sum of pair energies
  NVE (nve.cu:15-25)                 2-body oscillator closed form; time reversal;
                                     O(dt^2) energy error
  FIRE (fire.cu:13-46)               state-machine known answers; two disks relax
  Adam (base/kernels.cu:550-611)     first-step known answer
  rigid Adam (rigid_bumpy/kernels.cu: first-step known answer for centres, angles and
    1004-1093)                       vertex rings; rotation off keeps angles and rings
  stress (disk/kernels.cu:203-251)   2-particle closed form; virial == -dU/dA
  packing fraction (disk.cu:205-207) pi*sum r^2 / A
  cell reorder (particle.cu:1292-)   stable sort by cell id
  rigid geometry (rigid_bumpy)       ring radius, chord, I = m (R - r_v)^2, rotation
                                     round trip, torque balance
"""
import math

import numpy as np
import pytest

from _helpers import oracle, pkg, workloads

P = pkg()
W = workloads()


def orc_system(n, side, x, y, rad, mass=None, method=None, rc=None, vx=None, vy=None,
               kind=0, nv=0):
    s = P.System(kind, n, nv, backend=oracle())
    s.set_box(side, side)
    s.set_energy(1.0, 2.0)
    s.set_state(np.asarray(x, float), np.asarray(y, float), np.asarray(rad, float),
                np.ones(n) if mass is None else mass, vx, vy)
    if method is not None:
        if method == P.NBR_ALL:
            s.neighbor_config(P.NBR_ALL)
        else:
            spec = W.neighbor_spec(side, 2 * max(rad), (rc or 1.7) - 2 * max(rad), method)
            W.configure(s, spec)
        s.neighbor_init()
    return s


def forces(s):
    s.zero_forces()
    s.compute_forces()
    return s.download(P.F_FX), s.download(P.F_FY), s.download(P.F_PE)


# ---------------------------------------------------------------- pair law
def test_pair_law_closed_form():
    # sigma = 1.2, d = 1.0: overlap = 1/6, |F| = e_c*ov/sigma, U_pair = ov^2/2, each gets half
    s = orc_system(2, 10.0, [3.0, 4.0], [5.0, 5.0], [0.6, 0.6], method=P.NBR_ALL)
    fx, fy, pe = forces(s)
    ov = 1.0 - 1.0 / 1.2
    assert fx[0] == pytest.approx(-ov / 1.2, rel=1e-15)
    assert fx[1] == pytest.approx(ov / 1.2, rel=1e-15)
    assert fy[0] == 0.0 and fy[1] == 0.0
    assert pe[0] == pytest.approx(ov * ov / 4.0, rel=1e-15)
    assert pe.sum() == pytest.approx(ov * ov / 2.0, rel=1e-15)


def test_pair_law_zero_at_contact_and_beyond():
    for d in (1.2, 1.25, 1.6):
        s = orc_system(2, 10.0, [3.0, 3.0 + d], [5.0, 5.0], [0.6, 0.6], method=P.NBR_ALL)
        fx, fy, pe = forces(s)
        assert np.all(fx == 0) and np.all(pe == 0)


def test_pair_force_is_minus_energy_gradient():
    h = 1e-6
    def energy(d):
        s = orc_system(2, 10.0, [3.0, 3.0 + d * 0.6], [5.0, 5.0 + d * 0.8], [0.5, 0.7],
                       method=P.NBR_ALL)
        return forces(s)[2].sum()
    d0 = 1.05
    s = orc_system(2, 10.0, [3.0, 3.0 + d0 * 0.6], [5.0, 5.0 + d0 * 0.8], [0.5, 0.7],
                   method=P.NBR_ALL)
    fx, fy, _ = forces(s)
    dudr = (energy(d0 + h) - energy(d0 - h)) / (2 * h)
    # force on particle 1 along +r equals -dU/dr
    f_along = fx[1] * 0.6 + fy[1] * 0.8
    assert f_along == pytest.approx(-dudr, rel=1e-7)


def test_general_exponent():
    # n_c = 2.5: U_pair = e/n * ov^n, |F| = e ov^(n-1) / sigma
    s = orc_system(2, 10.0, [3.0, 4.0], [5.0, 5.0], [0.6, 0.6], method=P.NBR_ALL)
    s.set_energy(2.0, 2.5)
    fx, _, pe = forces(s)
    ov = 1.0 - 1.0 / 1.2
    assert fx[1] == pytest.approx(2.0 * ov ** 1.5 / 1.2, rel=1e-14)
    assert pe.sum() == pytest.approx(2.0 / 2.5 * ov ** 2.5, rel=1e-14)


def test_minimum_image_across_boundary():
    L = 10.0
    s = orc_system(2, L, [0.2, L - 0.6], [5.0, 5.0], [0.5, 0.5], method=P.NBR_ALL)
    fx, _, _ = forces(s)
    ov = 1.0 - 0.8 / 1.0
    assert fx[0] == pytest.approx(ov, rel=1e-15)   # pushed to +x, away from the image at -0.6
    assert fx[1] == pytest.approx(-ov, rel=1e-15)


def test_half_box_rounds_away_from_zero():
    # d = x1 - x2 = +L/2 exactly: round(0.5) = 1 (half away from zero) gives dx = -L/2,
    # so particle 1 is pushed towards -x; round-half-even would push it to +x.
    L = 2.0
    s = orc_system(2, L, [1.5, 0.5], [1.0, 1.0], [0.6, 0.6], method=P.NBR_ALL)
    fx, _, _ = forces(s)
    assert fx[0] < 0 and fx[1] > 0


# ---------------------------------------------------------------- neighbours
def test_cutoff_is_strict():
    # rc = 1.5 exactly; particles 1.5 apart are not neighbours, 1.4999 apart are
    for gap, expect in ((1.5, 0), (1.4999, 2)):
        s = P.System(0, 2, 0, backend=oracle())
        s.set_box(20.0, 20.0)
        s.set_state(np.array([5.0, 5.0 + gap]), np.array([5.0, 5.0]), np.array([0.5, 0.5]),
                    np.ones(2))
        s.neighbor_config(P.NBR_VERLET, 1.5, 1.0, 0, 0.0)
        s.neighbor_init()
        i, j = s.neighbor_pairs()
        assert len(i) == expect


def brute_pairs(x, y, L, rc):
    n = len(x)
    dx = x[:, None] - x[None, :]
    dy = y[:, None] - y[None, :]
    dx -= L * np.trunc(dx / L + np.copysign(0.5, dx))
    dy -= L * np.trunc(dy / L + np.copysign(0.5, dy))
    d2 = dx * dx + dy * dy
    amb = np.abs(d2 - rc * rc) < 1e-9
    inside = (d2 < rc * rc) & ~np.eye(n, dtype=bool)
    return inside, amb


@pytest.mark.parametrize("method", [2, 3])
def test_pair_sets_match_brute_force(method):
    st = W.disk_config(400, 0.8, seed=3)
    spec = W.neighbor_spec(st.side, method=method)
    s = P.System(0, st.n, 0, backend=oracle())
    s.set_box(st.side, st.side)
    s.set_state(st.x, st.y, st.rad, st.mass)
    W.configure(s, spec)
    s.neighbor_init()
    i, j = s.neighbor_pairs()
    got = np.zeros((st.n, st.n), bool)
    got[i, j] = True
    assert len(i) == got.sum()  # no duplicates
    want, amb = brute_pairs(st.x, st.y, st.side, spec.cutoff)
    assert np.array_equal(got & ~amb, want & ~amb)
    assert np.array_equal(got, got.T)  # full (both-direction) list


def test_cell_reorder_is_stable_sort_by_cell():
    st = W.disk_config(500, 0.7, seed=5)
    spec = W.neighbor_spec(st.side)
    s = W.make_disk_system(P.System, st, spec, backend=oracle())
    order = s.download(P.F_STATIC_INDEX)
    cs = st.side / spec.n_cells_dim
    n = spec.n_cells_dim
    cx = np.mod(np.floor(st.x / cs).astype(np.int64), n)
    cy = np.mod(np.floor(st.y / cs).astype(np.int64), n)
    cell = cx + cy * n
    assert np.array_equal(order, np.argsort(cell, kind="stable"))


def test_neighbor_capacity_grows_as_power_of_two():
    st = W.disk_config(300, 0.9, seed=9)
    s = W.make_disk_system(P.System, st, backend=oracle())
    stats = s.neighbor_stats()
    nn = s.download(P.F_NUM_NEIGHBORS)
    m = int(nn.max())
    assert stats["max_alloc"] == max(4, 2 ** math.ceil(math.log2(m)))


# ---------------------------------------------------------------- force sums
def test_newton_third_law_and_energy_sum():
    st = W.disk_config(600, 0.85, seed=11)
    s = W.make_disk_system(P.System, st, backend=oracle())
    fx, fy, pe = forces(s)
    fb = s.reduce(P.RED_FORCE_BALANCE)
    assert fb / st.n < 1e-14
    # independent brute-force energy
    dx = st.x[:, None] - st.x[None, :]
    dy = st.y[:, None] - st.y[None, :]
    L = st.side
    dx -= L * np.trunc(dx / L + np.copysign(0.5, dx))
    dy -= L * np.trunc(dy / L + np.copysign(0.5, dy))
    d = np.sqrt(dx * dx + dy * dy)
    sig = st.rad[:, None] + st.rad[None, :]
    ov = np.where((d < sig) & ~np.eye(st.n, dtype=bool), 1 - d / sig, 0.0)
    u = 0.5 * (ov ** 2).sum() / 2.0  # each unordered pair counted twice
    assert pe.sum() == pytest.approx(u, rel=1e-12)
    with np.errstate(divide="ignore", invalid="ignore"):
        fm = np.where(ov > 0, ov / (sig * d), 0.0)
    np.testing.assert_allclose(fx, (fm * dx).sum(axis=1), rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(fy, (fm * dy).sum(axis=1), rtol=1e-12, atol=1e-15)


# ---------------------------------------------------------------- NVE
def test_two_body_oscillator():
    sigma, delta0 = 1.2, 0.05
    s = orc_system(2, 10.0, [3.0, 3.0 + sigma - delta0], [5.0, 5.0], [0.6, 0.6],
                   method=P.NBR_ALL)
    forces(s)
    omega = math.sqrt(2.0) / sigma
    dt = 1e-4
    n = int(round((math.pi / 4) / omega / dt))
    s.nve_run(dt, n)
    x = s.download(P.F_X)
    t = n * dt
    expect = sigma - delta0 * math.cos(omega * t)
    assert (x[1] - x[0]) == pytest.approx(expect, abs=1e-8)


def test_time_reversibility():
    st = W.disk_config(256, 0.82, seed=21)
    s = W.make_disk_system(P.System, st, backend=oracle())
    vx, vy = W.gaussian_velocities(st.n, 1e-3, 1)
    s.upload(P.F_VX, vx)
    s.upload(P.F_VY, vy)
    forces(s)
    x0, y0 = s.download(P.F_X), s.download(P.F_Y)
    s.nve_run(0.005, 200)
    s.upload(P.F_VX, -s.download(P.F_VX))
    s.upload(P.F_VY, -s.download(P.F_VY))
    s.nve_run(0.005, 200)
    np.testing.assert_allclose(s.download(P.F_X), x0, atol=1e-10)
    np.testing.assert_allclose(s.download(P.F_Y), y0, atol=1e-10)


def energy_fluct(dt, steps):
    st = W.disk_config(128, 0.84, seed=4)
    s = W.make_disk_system(P.System, st, backend=oracle())
    s.minimize_fire(P.fire_params(max_steps=200))
    vx, vy = W.gaussian_velocities(st.n, 1e-3, 2)
    s.upload(P.F_VX, vx)
    s.upload(P.F_VY, vy)
    forces(s)
    es = []
    for _ in range(steps):
        s.nve_run(dt, int(round(0.05 / dt)))
        es.append(s.reduce(P.RED_PE) + s.reduce(P.RED_KE))
    es = np.array(es)
    return es.max() - es.min()


def test_energy_error_scales_as_dt_squared():
    a = energy_fluct(0.02, 20)
    b = energy_fluct(0.01, 20)
    assert 2.5 < a / b < 6.0


# ---------------------------------------------------------------- FIRE
def test_fire_state_machine_known_answers():
    s = orc_system(2, 10.0, [3.0, 4.0], [5.0, 5.0], [0.6, 0.6], method=P.NBR_ALL)
    s.zero_velocities()
    forces(s)
    st = s.fire_state(0.1, 0.01)
    s.fire_step(st)
    # from rest the power is 0 -> uphill branch: dt halves, alpha resets, N_bad = 1
    assert st.power == 0.0 and st.n_bad == 1 and st.n_good == 0
    assert st.dt == pytest.approx(0.005) and st.alpha == pytest.approx(0.1)
    for k in range(5):
        s.fire_step(st)
        assert st.power > 0 and st.n_good == k + 1 and st.dt == pytest.approx(0.005)
    s.fire_step(st)  # N_good = 6 > N_min = 5: dt *= 1.1, alpha *= 0.99
    assert st.n_good == 6
    assert st.dt == pytest.approx(0.0055) and st.alpha == pytest.approx(0.099)


def test_fire_uphill_back_step_known_answer():
    # velocities against the forces -> P < 0: dt is halved first, then x -= v*dt_new/2,
    # v = 0, kick(dt_new/2), mix (v parallel to f: unchanged), drift(dt_new)
    s = orc_system(2, 10.0, [3.0, 4.0], [5.0, 5.0], [0.6, 0.6], method=P.NBR_ALL,
                   vx=np.array([0.3, -0.3]), vy=np.zeros(2))
    fx, _, _ = forces(s)
    x0 = s.download(P.F_X)
    st = s.fire_state(0.1, 0.01)
    s.fire_step(st)
    assert st.power < 0 and st.n_bad == 1
    dt = 0.005
    v0 = np.array([0.3, -0.3])
    expect = x0 - v0 * dt / 2.0 + fx * (dt / 2.0) * dt
    np.testing.assert_allclose(s.download(P.F_X), expect, rtol=0, atol=1e-15)


def test_fire_relaxes_two_disks_to_contact():
    s = orc_system(2, 10.0, [3.0, 4.0], [5.0, 5.0], [0.6, 0.6], method=P.NBR_ALL)
    r = s.minimize_fire(P.fire_params(avg_pe_target=1e-20, avg_pe_diff_target=0.0,
                                      max_steps=100000))
    x = s.download(P.F_X)
    assert r.converged == 1
    assert (x[1] - x[0]) >= 1.2 - 1e-9
    assert s.reduce(P.RED_PE) < 2e-20


def test_fire_three_disk_chain():
    s = orc_system(3, 10.0, [3.0, 3.9, 4.8], [5.0, 5.0, 5.0], [0.5, 0.5, 0.5],
                   method=P.NBR_ALL)
    s.minimize_fire(P.fire_params(avg_pe_target=1e-20, avg_pe_diff_target=0.0))
    x = np.sort(s.download(P.F_X))
    assert np.all(np.diff(x) >= 1.0 - 1e-9)
    assert abs(s.download(P.F_Y) - 5.0).max() < 1e-12  # collinear forces keep y


def test_fire_max_force_criterion():
    st = W.disk_config(200, 0.86, seed=8)
    s = W.make_disk_system(P.System, st, backend=oracle())
    r = s.minimize_fire(P.fire_params(avg_pe_target=0.0, avg_pe_diff_target=0.0,
                                      fmax_target=1e-10, max_steps=200000))
    assert r.converged == 1
    fx, fy = s.download(P.F_FX), s.download(P.F_FY)
    assert np.max(fx * fx + fy * fy) < 1e-20


# ---------------------------------------------------------------- Adam
def test_adam_first_step_known_answer():
    s = orc_system(2, 10.0, [3.0, 4.0], [5.0, 5.0], [0.6, 0.6], method=P.NBR_ALL)
    x0 = s.download(P.F_X)
    alpha, eps = 1e-3, 1e-8
    s.adam_step(0, alpha, 0.9, 0.999, eps)
    fx = s.download(P.F_FX)
    # m_hat = -f, v_hat = f^2 on the first step: dx = alpha * f / (|f| + eps)
    np.testing.assert_allclose(s.download(P.F_X) - x0, alpha * fx / (np.abs(fx) + eps),
                               rtol=1e-10)
    np.testing.assert_allclose(s.download(P.F_Y), 5.0, rtol=0, atol=0)


# ---------------------------------------------------------------- stress
def test_pressure_two_particles_closed_form():
    L = 10.0
    s = orc_system(2, L, [3.0, 4.0], [5.0, 5.0], [0.6, 0.6], method=P.NBR_ALL)
    p = s.reduce(P.RED_PRESSURE)
    ov = 1.0 - 1.0 / 1.2
    f = ov / 1.2
    assert p == pytest.approx(1.0 * f / (2 * L * L), rel=1e-14)


def test_virial_pressure_equals_minus_dU_dA():
    st = W.disk_config(300, 0.9, seed=13)
    s = W.make_disk_system(P.System, st, backend=oracle())
    s.minimize_fire(P.fire_params(max_steps=300))
    p = s.reduce(P.RED_PRESSURE)
    x, y = s.download(P.F_X), s.download(P.F_Y)
    L = s.box()[0]
    eps = 1e-6

    def u_at(f):
        t = orc_system(st.n, L * f, x * f, y * f, st.rad, method=3)
        return forces(t)[2].sum()

    dA = (L * (1 + eps)) ** 2 - (L * (1 - eps)) ** 2
    p_fd = -(u_at(1 + eps) - u_at(1 - eps)) / dA
    assert p == pytest.approx(p_fd, rel=1e-5)


def test_packing_fraction():
    st = W.disk_config(300, 0.77, seed=1)
    s = W.make_disk_system(P.System, st, backend=oracle())
    assert s.reduce(P.RED_PACKING_FRACTION) == pytest.approx(0.77, rel=1e-13)
    s.scale_to_packing_fraction(0.8)
    assert s.reduce(P.RED_PACKING_FRACTION) == pytest.approx(0.8, rel=1e-13)


def test_canonical_sum_accuracy():
    rng = np.random.default_rng(0)
    lib = oracle().lib
    f = lib.orc_canonical_sum
    import ctypes as C
    f.restype = C.c_double
    f.argtypes = [C.c_void_p, C.c_int64]
    for n in (1, 63, 64, 65, 4096, 100001):
        a = rng.integers(-1000, 1000, n).astype(np.float64)
        assert f(a.ctypes.data, n) == float(a.sum())  # exact for integers
        b = rng.standard_normal(n)
        assert f(b.ctypes.data, n) == pytest.approx(math.fsum(b), rel=1e-12, abs=1e-12)


# ---------------------------------------------------------------- rigid bumpy
def bumpy_small(n=2, rotation=True, x=None, y=None, angle=None, side=20.0):
    rad = np.array([0.7, 0.5][:n] if n <= 2 else W.bidisperse_radii(n))
    nverts = np.where(rad == 0.7, 28, 20).astype(np.int64)
    rv = W.vertex_radius()
    s = P.System(1, n, int(nverts.sum()), backend=oracle())
    s.set_box(side, side)
    s.set_energy(1.0, 2.0)
    s.set_state(np.asarray(x, float), np.asarray(y, float), rad, np.ones(n))
    s.upload(P.F_ANGLE, np.asarray(angle, float))
    s.bumpy_setup(nverts, rv, rotation)
    return s, rad, nverts, rv


def test_bumpy_ring_geometry_and_inertia():
    s, rad, nverts, rv = bumpy_small(2, x=[5.0, 12.0], y=[5.0, 6.0], angle=[0.3, 1.1])
    vx, vy = s.download(P.F_VERTEX_X), s.download(P.F_VERTEX_Y)
    vm = s.download(P.F_VERTEX_MASS)
    moi = s.download(P.F_MOI)
    start = 0
    for p, (cx, cy) in enumerate([(5.0, 5.0), (12.0, 6.0)]):
        n = nverts[p]
        px, py = vx[start:start + n], vy[start:start + n]
        R = rad[p] - rv
        np.testing.assert_allclose(np.hypot(px - cx, py - cy), R, rtol=1e-14)
        chord = np.hypot(np.diff(px), np.diff(py))
        np.testing.assert_allclose(chord, 2 * R * math.sin(math.pi / n), rtol=1e-12)
        assert vm[start:start + n].sum() == pytest.approx(1.0, rel=1e-14)
        assert moi[p] == pytest.approx(R * R, rel=1e-13)
        start += n
    assert rv == pytest.approx(0.0676365, rel=1e-6)  # value quoted for cfg 3


def test_bumpy_rotation_round_trip():
    s, rad, nverts, rv = bumpy_small(2, x=[5.0, 12.0], y=[5.0, 6.0], angle=[0.3, 1.1])
    vx0, vy0 = s.download(P.F_VERTEX_X), s.download(P.F_VERTEX_Y)
    s.upload(P.F_OMEGA, np.array([3.0, -2.0]))
    s.drift(0.1)
    th = s.download(P.F_ANGLE)
    assert th[0] == pytest.approx(0.6) and th[1] == pytest.approx(0.9)
    s.upload(P.F_OMEGA, np.array([-3.0, 2.0]))
    s.drift(0.1)
    np.testing.assert_allclose(s.download(P.F_VERTEX_X), vx0, atol=1e-13)
    np.testing.assert_allclose(s.download(P.F_VERTEX_Y), vy0, atol=1e-13)


def test_bumpy_torque_balance_and_contact():
    # two rings close enough that a few vertices overlap
    s, rad, nverts, rv = bumpy_small(2, x=[5.0, 6.15], y=[5.0, 5.05], angle=[0.0, 0.1])
    s.neighbor_config(P.NBR_ALL)
    s.bumpy_neighbor_config(1.5 * 2 * rv, 1.4)
    s.neighbor_init()
    fx, fy, pe = forces(s)
    tq = s.download(P.F_TORQUE)
    assert pe.sum() > 0
    assert abs(fx.sum()) < 1e-15 and abs(fy.sum()) < 1e-15
    # angular momentum balance about the origin: sum(tau_i + x_i x F_i) = 0
    x, y = np.array([5.0, 6.15]), np.array([5.0, 5.05])
    total = (tq + x * fy - y * fx).sum()
    assert abs(total) < 1e-13


@pytest.mark.parametrize("rotation", [True, False])
def test_bumpy_adam_first_step_known_answer(rotation):
    # two overlapping rings: on the first step m_hat = -f, v_hat = f^2 for the force and
    # the torque alike, so the centre moves by alpha f / (|f| + eps), the angle by
    # alpha tau / (|tau| + eps) (rotation on) and every vertex keeps its place in the
    # ring, turned by that angle about the moved centre
    x0, y0, a0 = np.array([5.0, 6.15]), np.array([5.0, 5.05]), np.array([0.0, 0.1])
    s, rad, nverts, rv = bumpy_small(2, rotation=rotation, x=x0, y=y0, angle=a0)
    s.neighbor_config(P.NBR_ALL)
    s.bumpy_neighbor_config(1.5 * 2 * rv, 1.4)
    s.neighbor_init()
    vx0, vy0 = s.download(P.F_VERTEX_X), s.download(P.F_VERTEX_Y)
    alpha, eps = 1e-3, 1e-8
    s.adam_reset()
    s.adam_step(0, alpha, 0.9, 0.999, eps)
    fx, fy, tq = s.download(P.F_FX), s.download(P.F_FY), s.download(P.F_TORQUE)
    assert np.all(np.abs(fx) > 1e-6) and np.all(np.abs(tq) > 1e-9)
    x1, y1, a1 = s.download(P.F_X), s.download(P.F_Y), s.download(P.F_ANGLE)
    np.testing.assert_allclose(x1 - x0, alpha * fx / (np.abs(fx) + eps), rtol=1e-10)
    np.testing.assert_allclose(y1 - y0, alpha * fy / (np.abs(fy) + eps), rtol=1e-10)
    da = alpha * tq / (np.abs(tq) + eps) if rotation else np.zeros(2)
    np.testing.assert_allclose(a1 - a0, da, rtol=1e-10, atol=0)
    owner = np.repeat(np.arange(2), nverts)
    rx, ry = vx0 - x0[owner], vy0 - y0[owner]
    c, sn = np.cos(da[owner]), np.sin(da[owner])
    np.testing.assert_allclose(s.download(P.F_VERTEX_X), x1[owner] + rx * c - ry * sn,
                               rtol=0, atol=1e-13)
    np.testing.assert_allclose(s.download(P.F_VERTEX_Y), y1[owner] + rx * sn + ry * c,
                               rtol=0, atol=1e-13)


def test_bumpy_adam_relaxes_overlaps():
    # Adam from overlapping rings lowers the energy; the rings stay rigid
    x0, y0 = np.array([5.0, 6.15]), np.array([5.0, 5.05])
    s, rad, nverts, rv = bumpy_small(2, x=x0, y=y0, angle=[0.0, 0.1])
    s.neighbor_config(P.NBR_ALL)
    s.bumpy_neighbor_config(1.5 * 2 * rv, 1.4)
    s.neighbor_init()
    s.zero_forces()
    s.compute_forces()
    pe0 = s.reduce(P.RED_PE)
    steps = s.minimize_adam(alpha=1e-3, max_steps=2000)
    assert 0 < steps <= 2000
    assert s.reduce(P.RED_PE) < 1e-3 * pe0
    x, y = s.download(P.F_X), s.download(P.F_Y)
    owner = np.repeat(np.arange(2), nverts)
    d = np.hypot(s.download(P.F_VERTEX_X) - x[owner], s.download(P.F_VERTEX_Y) - y[owner])
    ring = (rad - rv)[owner]
    np.testing.assert_allclose(d, ring, rtol=0, atol=1e-12)


# ---------------------------------------------------------------- shared sin/cos
def _sincos(x):
    import ctypes as C

    from _helpers import ORACLE_SO
    lib = C.CDLL(ORACLE_SO)
    x = np.ascontiguousarray(x, dtype=np.float64)
    s, c = np.empty_like(x), np.empty_like(x)
    ptr = lambda a: a.ctypes.data_as(C.c_void_p)
    lib.orc_sincos.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64]
    assert lib.orc_sincos(ptr(x), ptr(s), ptr(c), x.size) == 0
    return s, c


@pytest.mark.parametrize("lim,max_ulp", [(math.pi / 4, 1.0), (7.0, 2.0), (100.0, 2.0),
                                         (1e5, 3.0)])
def test_shared_sincos_against_libm(lim, max_ulp):
    # dpm_sincos (rigid-body rotations on both sides) vs the C library's correctly
    # rounded-to-0.5-ulp sin/cos: together at most max_ulp apart
    rng = np.random.default_rng(7)
    x = np.concatenate([rng.uniform(-lim, lim, 200000),
                        [0.0, -0.0, 1e-300, 5e-324, math.pi / 4, math.pi / 2, math.pi,
                         -math.pi / 2, 3 * math.pi / 2, 2 * math.pi]])
    x = x[np.abs(x) <= lim]
    s, c = _sincos(x)
    for got, ref in ((s, np.sin(x)), (c, np.cos(x))):
        err = np.abs(got - ref) / np.spacing(np.maximum(np.abs(ref), 2.0 ** -1022))
        assert err.max() <= max_ulp, (lim, err.max())
    assert np.all(s * s + c * c == pytest.approx(1.0, abs=4e-16))
    # exact special values and symmetries of the reduction
    s0, c0 = _sincos(np.array([0.0, -0.0]))
    assert s0[0] == 0.0 and c0[0] == 1.0 and math.copysign(1, s0[1]) == -1.0
    sp, cp = _sincos(x)
    sn, cn = _sincos(-x)
    assert np.array_equal(sn, -sp) and np.array_equal(cn, cp)
