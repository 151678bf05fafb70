"""GPU tests of the batched jamming engine (dpm_batch_*): every system's result is
bit-identical to the same system jammed alone by dpm_jam (Verlet method, same cutoff and
threshold), whatever the other systems in the batch are, and to the oracle's jam of that
system (orc_jam, the CPU restatement of scripts/runArthurJamWithFire.cpp:64-106)."""
import numpy as np
import pytest

from _helpers import oracle, pkg, workloads

P = pkg()
W = workloads()
pytestmark = pytest.mark.gpu

FIELDS = ("status", "compression_steps", "fire_steps_total", "pressure", "packing_fraction",
          "pe", "box_length", "contacts")


def say(*a):
    print("[gpu-case]", *a, flush=True)


def params(max_steps=1500, fmax=1e-10, strain=5e-3, csteps=400):
    return P.jam_params(pressure_target=1e-3, window_rel=0.2, strain_step=strain,
                        max_compression_steps=csteps,
                        fire=P.fire_params(avg_pe_target=0.0, avg_pe_diff_target=0.0,
                                           fmax_target=fmax, max_steps=max_steps))


def systems(spec):
    return [W.disk_config(n, phi, seed=seed) for n, phi, seed in spec]


def plain(st, prm, backend=None):
    s = W.make_disk_system(P.System, st, W.neighbor_spec(st.side, method=2), backend=backend)
    r = s.jam(prm)
    x, y = s.download(P.F_X), s.download(P.F_Y)
    return r, x, y


def same_result(got, ref, k):
    for f in FIELDS:
        assert getattr(got, f) == getattr(ref, f), (k, f, getattr(got, f), getattr(ref, f))


def batched(sts, prm, max_batched_steps=10**8):
    b = P.Batch([st.n for st in sts])
    b.config(1.0, 2.0, 1.7, 3.0 * 0.15 * 0.15)
    for k, st in enumerate(sts):
        b.set_system(k, st.x, st.y, st.rad, st.mass, st.side)
    res, steps = b.jam(prm, max_batched_steps)
    pos = [b.get_system(k) for k in range(len(sts))]
    b.close()
    return res, pos, steps


SPEC = [(64, 0.80, 1), (100, 0.82, 2), (150, 0.78, 3), (257, 0.80, 4), (70, 0.83, 5)]


def test_batched_equals_each_system_alone():
    sts = systems(SPEC)
    prm = params()
    res, pos, steps = batched(sts, prm)
    say("batched steps", steps)
    for k, st in enumerate(sts):
        r, x, y = plain(st, prm)
        ro, ox, oy = plain(st, prm, backend=oracle())
        say("system", k, st.n, r.status, r.compression_steps, r.fire_steps_total)
        same_result(res[k], r, k)
        same_result(res[k], ro, k)
        bx, by, side = pos[k]
        assert np.array_equal(bx, x) and np.array_equal(by, y), k
        assert np.array_equal(bx, ox) and np.array_equal(by, oy), k
        assert side == r.box_length == ro.box_length


def test_batched_minimization_equals_minimize_fire_alone():
    """dpm_batch_minimize: minimizeFire of every system at once; each system's steps,
    energy, pressure, contacts, packing fraction and positions equal dpm_minimize_fire on
    that system alone and the oracle's, bit for bit."""
    sts = systems(SPEC + [(2048, 0.84, 6)])
    fire = P.fire_params(avg_pe_target=0.0, avg_pe_diff_target=0.0, fmax_target=1e-9,
                         max_steps=4000)
    b = P.Batch([st.n for st in sts])
    b.config(1.0, 2.0, 1.7, 3.0 * 0.15 * 0.15)
    for k, st in enumerate(sts):
        b.set_system(k, st.x, st.y, st.rad, st.mass, st.side)
    res, steps = b.minimize(fire)
    say("batched minimisation steps", steps)
    for k, st in enumerate(sts):
        for be in (None, oracle()):
            s = W.make_disk_system(P.System, st, W.neighbor_spec(st.side, method=2), backend=be)
            fr = s.minimize_fire(fire)
            assert res[k].fire_steps_total == fr.steps, (k, be)
            assert res[k].status == (0 if fr.converged else 1), k
            assert res[k].pe == s.reduce(P.RED_PE), k
            assert res[k].pressure == s.reduce(P.RED_PRESSURE), k
            assert res[k].contacts == s.reduce(P.RED_CONTACTS), k
            assert res[k].packing_fraction == s.reduce(P.RED_PACKING_FRACTION), k
            assert res[k].box_length == st.side and res[k].compression_steps == 0
            bx, by, _ = b.get_system(k)
            assert np.array_equal(bx, s.download(P.F_X)), k
            assert np.array_equal(by, s.download(P.F_Y)), k
    b.close()


def test_4096_uneven_systems_batched_jamming():
    """The batched-jamming configuration: 4,096 seeded systems, N log-uniform in
    [64, 2048], phi0 = 0.80, jammed in one batch with the relaxed protocol the benchmark
    states (strain 1e-2, FIRE to |F|max < 1e-6 in at most 3,000 steps, target pressure
    1e-2 within a window of 0.5).  Every system must end jammed (status 0: pressure inside
    the window with contacts), and the 16 smallest systems equal the oracle's jam bit for
    bit."""
    sts = W.batched_systems(4096, 12345)
    prm = P.jam_params(pressure_target=1e-2, window_rel=0.5, strain_step=1e-2,
                       max_compression_steps=100000,
                       fire=P.fire_params(avg_pe_target=0.0, avg_pe_diff_target=0.0,
                                          fmax_target=1e-6, max_steps=3000))
    res, pos, steps = batched(sts, prm)
    status = np.array([r.status for r in res])
    say("4096 systems: batched steps", steps, "statuses", np.unique(status, return_counts=True))
    assert np.all(status == 0)
    for r in res:
        assert 1e-2 <= r.pressure <= 1.5e-2 and r.contacts > 0
    small = np.argsort([st.n for st in sts], kind="stable")[:16]
    for k in small:
        ro, ox, oy = plain(sts[k], prm, backend=oracle())
        say("4096 systems: oracle sample", k, sts[k].n, ro.compression_steps, ro.fire_steps_total)
        same_result(res[k], ro, int(k))
        bx, by, side = pos[k]
        assert np.array_equal(bx, ox) and np.array_equal(by, oy), k


def test_large_systems_cell_built_rows():
    """Systems of 128 particles and more get their Verlet rows from a cell grid; the rows,
    and so every result bit, equal the all-pairs rows of the plain path.  The largest
    system fills the grid's particle capacity (16,384)."""
    sts = systems([(2048, 0.80, 11), (4500, 0.82, 12), (300, 0.80, 13), (16384, 0.80, 14)])
    prm = params(max_steps=300, fmax=1e-8, strain=1e-2, csteps=40)
    res, pos, steps = batched(sts, prm)
    say("large batched steps", steps)
    for k, st in enumerate(sts):
        r, x, y = plain(st, prm)
        say("system", k, st.n, r.status, r.compression_steps, r.fire_steps_total)
        for f in FIELDS:
            assert getattr(res[k], f) == getattr(r, f), (k, f, getattr(res[k], f), getattr(r, f))
        bx, by, side = pos[k]
        assert np.array_equal(bx, x) and np.array_equal(by, y), k


def test_box_smaller_than_twice_the_cutoff():
    # 6 disks: half the box is under the cutoff, so the batched force pass takes the
    # reference's minimum image (round half away) instead of the branch-free one
    sts = systems([(6, 0.80, 41), (64, 0.80, 42)])
    assert sts[0].side < 2 * 1.7
    prm = params(max_steps=800)
    res, pos, steps = batched(sts, prm)
    for k, st in enumerate(sts):
        r, x, y = plain(st, prm)
        say("system", k, st.n, st.side, r.status, r.fire_steps_total)
        for f in FIELDS:
            assert getattr(res[k], f) == getattr(r, f), (k, f, getattr(res[k], f), getattr(r, f))
        bx, by, side = pos[k]
        assert np.array_equal(bx, x) and np.array_equal(by, y), k


def test_result_same_alone_and_in_any_mix():
    sts = systems(SPEC)
    prm = params(max_steps=800)
    mix, _, _ = batched(sts, prm)
    alone, _, _ = batched([sts[2]], prm)
    rev, _, _ = batched(list(reversed(sts)), prm)
    for f in FIELDS:
        assert getattr(mix[2], f) == getattr(alone[0], f) == getattr(rev[2], f), f


def test_unfinished_systems_report_running():
    sts = systems(SPEC[:2])
    res, _, steps = batched(sts, params(), max_batched_steps=64)
    assert steps == 64
    assert all(r.status == -1 for r in res)


def test_bad_arguments_fail_loudly():
    with pytest.raises(P.DpmError):
        P.Batch([64, 0])
    b = P.Batch([10])
    with pytest.raises(P.DpmError):
        b.jam(params())  # no particles set
    with pytest.raises(P.DpmError):
        b.set_system(3, np.zeros(10), np.zeros(10), np.ones(10), np.ones(10), 5.0)


def test_autotune_keeps_results():
    sts = systems(SPEC[:3])
    prm = params(max_steps=600)
    ref, _, _ = batched(sts, prm)
    b = P.Batch([st.n for st in sts])
    b.config(1.0, 2.0, 1.7, 3.0 * 0.15 * 0.15)
    for k, st in enumerate(sts):
        b.set_system(k, st.x, st.y, st.rad, st.mass, st.side)
    cut, chunk, ms = b.autotune(prm, trial_steps=64, cutoffs=(1.6, 1.9), chunks=(8, 32))
    say("autotune", cut, chunk, ms)
    assert cut in (1.6, 1.9) and chunk in (8, 32) and ms > 0
    blk, pct, ms = b.autotune_shape(prm, trial_steps=64)
    say("autotune_shape", blk, pct, ms)
    assert blk in (64, 128, 256) and pct in (100, 125, 150) and ms > 0
    res, _ = b.jam(prm)
    b.close()
    for k in range(len(sts)):
        for f in FIELDS:
            assert getattr(res[k], f) == getattr(ref[k], f), (k, f)


@pytest.mark.parametrize("block,pct", [(64, 100), (128, 150), (256, 400)])
def test_launch_shape_keeps_results(block, pct):
    """The force pass's workgroup size and the cell width leave every result bit alone."""
    sts = systems(SPEC[:4])
    prm = params(max_steps=600)
    ref, _, _ = batched(sts, prm)
    b = P.Batch([st.n for st in sts])
    b.config(1.0, 2.0, 1.7, 3.0 * 0.15 * 0.15)
    for k, st in enumerate(sts):
        b.set_system(k, st.x, st.y, st.rad, st.mass, st.side)
    b.set_option(P.OPT_FORCE_BLOCK, block)
    b.set_option(P.OPT_CELL_PCT, pct)
    for opt, bad in ((P.OPT_FORCE_BLOCK, 96), (P.OPT_CELL_PCT, 99), (P.OPT_CELL_PCT, 401)):
        with pytest.raises(P.DpmError):
            b.set_option(opt, bad)
    res, _ = b.jam(prm)
    b.close()
    say("shape", block, pct)
    for k in range(len(sts)):
        for f in FIELDS:
            assert getattr(res[k], f) == getattr(ref[k], f), (k, f)
