"""GPU tests of the batched jamming engine (dpm_batch_*): every system's result is
bit-identical to the same system jammed alone by dpm_jam (Verlet method, same cutoff and
threshold), whatever the other systems in the batch are."""
import numpy as np
import pytest

from _helpers import pkg, workloads

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


def plain(st, prm):
    s = W.make_disk_system(P.System, st, W.neighbor_spec(st.side, method=2))
    r = s.jam(prm)
    x, y = s.download(P.F_X), s.download(P.F_Y)
    return r, x, y


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
        say("system", k, st.n, r.status, r.compression_steps, r.fire_steps_total)
        for f in FIELDS:
            assert getattr(res[k], f) == getattr(r, f), (k, f, getattr(res[k], f), getattr(r, f))
        bx, by, side = pos[k]
        assert np.array_equal(bx, x) and np.array_equal(by, y), k
        assert side == r.box_length


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
    res, _ = b.jam(prm)
    b.close()
    for k in range(len(sts)):
        for f in FIELDS:
            assert getattr(res[k], f) == getattr(ref[k], f), (k, f)
