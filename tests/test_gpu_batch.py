"""Batched stepping (ff_batch_*): every member's results equal those of the
same domain stepped alone with ff_step, bit for bit (moments, per-step FP64
wrench, vertex forces), whatever the tile shape, the membership changes and
the mix of domains with and without a boundary."""
import sys

import numpy as np
import pytest

from freeflow_amd import capi, workloads as wl

pytestmark = pytest.mark.gpu


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def scenes(seed, n=3, nx=(96, 220), ny=(64, 120), seg=(60, 160)):
    return wl.batched_scenes(n, seed, nx, ny, seg)


def make(scene, storage="f32", history=64):
    d = capi.Domain(**scene.domain_args(), storage=storage, stage_fp16=True)
    scene.setup(d, history=history)
    return d


def same(a, b, what):
    ma, mb = a.moments(0), b.moments(0)
    assert np.array_equal(ma, mb), f"{what}: {np.sum(ma != mb)} moments differ"
    assert np.array_equal(a.moments(1), b.moments(1)), f"{what}: other buffer differs"


def test_batch_equals_individual_steps_2d():
    sc = scenes(11)
    note(f"case batch2d sizes {[(s.nx, s.ny, len(s.swimmer.elems)) for s in sc]}")
    alone = [make(s) for s in sc]
    mixed = [make(s) for s in sc]
    b = capi.Batch(2, "f32", stage_fp16=True)
    for d in mixed:
        b.add(d)
    for d in alone:
        d.step(25)
    b.step(25)
    b.sync()
    for i, (a, m) in enumerate(zip(alone, mixed)):
        same(a, m, f"domain {i}")
        assert np.array_equal(a.body_wrench(), m.body_wrench())
        assert np.array_equal(a.wrench_history(25), m.wrench_history(25))
        assert np.array_equal(a.vertex_forces(), m.vertex_forces())
        assert np.array_equal(a.boundary()[0], m.boundary()[0])
    # membership change: drop the middle swimmer, add a boundary-free vortex
    # street; everybody keeps matching the solo runs
    b.remove(mixed[1])
    nx, ny = 120, 60
    flags = wl.vortex_street_flags(nx, ny)
    m0 = wl.noisy_inflow_moments(flags, nx, ny, 0.05, 1e-3, seed=5)
    vs_alone = capi.Domain(2, nx, ny, viscosity=0.005)
    vs_mixed = capi.Domain(2, nx, ny, viscosity=0.005)
    for d in (vs_alone, vs_mixed):
        d.set_flags(flags)
        d.set_moments(m0)
    b.add(vs_mixed)
    for d in (alone[0], alone[2], vs_alone):
        d.step(13)
    mixed[1].step(13)  # detached again: steps on its own
    alone[1].step(13)
    b.step(13)
    b.sync()
    for a, m, what in ((alone[0], mixed[0], "d0"), (alone[2], mixed[2], "d2"), (vs_alone, vs_mixed, "street"),
                       (alone[1], mixed[1], "detached")):
        same(a, m, what)
    with pytest.raises(capi.FreeFlowError):
        mixed[0].step(1)


def test_graph_replays_between_eager_work():
    """Runs of 8 or more steps replay an 8-step graph: replay it several
    times with eager work (host reads, single-step replays, a wave change)
    between the replays, checking every stop against the domains alone."""
    sc = scenes(21, n=2)
    note(f"case batch replays sizes {[(s.nx, s.ny) for s in sc]}")
    alone = [make(s) for s in sc]
    mixed = [make(s) for s in sc]
    b = capi.Batch(2, "f32", stage_fp16=True)
    for d in mixed:
        b.add(d)
    for k, n in enumerate((8, 3, 17, 8, 1)):
        if k == 3:  # an action between replays
            for d in alone + mixed:
                p = sc[0].swimmer.params.copy() if d in (alone[0], mixed[0]) else sc[1].swimmer.params.copy()
                p[0] *= 0.5
                d.set_wave_params(p[None, :])
        for d in alone:
            d.step(n)
        b.step(n)
        b.sync()
        for i, (a, m) in enumerate(zip(alone, mixed)):
            same(a, m, f"stop {k} domain {i}")
            assert np.array_equal(a.body_wrench(), m.body_wrench())


@pytest.mark.parametrize("tile", [1, 2, 3, 4])
def test_tile_shapes_give_identical_results(tile):
    sc = scenes(21, n=2)
    note(f"case tile {tile}")
    ref = [make(s, history=0) for s in sc]
    got = [make(s, history=0) for s in sc]
    b0 = capi.Batch(2, "f32")
    b0.set_tile(0)
    bt = capi.Batch(2, "f32")
    bt.set_tile(tile)
    for d in ref:
        b0.add(d)
    for d in got:
        bt.add(d)
    b0.step(20)
    bt.step(20)
    assert bt.tile()[0] == tile
    for a, m in zip(ref, got):
        same(a, m, f"tile {tile}")
        assert np.array_equal(a.body_wrench(), m.body_wrench())


@pytest.mark.parametrize("dim,tile", [(2, t) for t in range(5)] + [(3, t) for t in range(7)])
def test_members_smaller_than_a_tile(dim, tile):
    # members of a few cells next to a regular one, in every tile shape:
    # each equals the same domain stepped alone
    shapes = [(5, 3), (17, 9), (64, 40)] if dim == 2 else [(5, 4, 3), (19, 9, 7), (40, 24, 20)]
    note(f"case tiny members dim={dim} tile={tile} {shapes}")
    rng = np.random.default_rng(50 + tile)

    def build(shape, seed):
        n = int(np.prod(shape))
        u = np.random.default_rng(seed).uniform(-0.03, 0.03, (n, dim))
        d = capi.Domain(dim, *shape, viscosity=0.02)
        d.set_moments(wl.equilibrium_moments(u))
        return d

    seeds = rng.integers(1 << 30, size=len(shapes))
    alone = [build(sh, sd) for sh, sd in zip(shapes, seeds)]
    mixed = [build(sh, sd) for sh, sd in zip(shapes, seeds)]
    b = capi.Batch(dim, "f32")
    b.set_tile(tile)
    for d in mixed:
        b.add(d)
    for d in alone:
        d.set_tile(tile)
        d.step(12)
    b.step(12)
    b.sync()
    for i, (a, m) in enumerate(zip(alone, mixed)):
        same(a, m, f"member {i} {shapes[i]}")


def test_autotune_and_single_domain_tiles():
    s = scenes(31, n=1)[0]
    note("case autotune")
    a, m = make(s, history=0), make(s, history=0)
    t = m.autotune()
    assert 0 <= t < 5
    a.step(15)
    m.step(15)
    same(a, m, f"autotuned tile {t}")
    b = capi.Batch(2, "f32")
    c = make(s, history=0)
    b.add(c)
    b.step(15)
    tile, tx, ty = b.tile()
    assert 0 <= tile < 5 and tx in (32, 64) and ty in (8, 16, 32)
    same(a, c, "batch autotune")


@pytest.mark.parametrize("storage", ["f16", "bf16", "f16u"])
def test_batch_16bit_storage(storage):
    sc = scenes(41, n=2)
    note(f"case batch storage {storage}")
    alone = [make(s, storage=storage) for s in sc]
    mixed = [make(s, storage=storage) for s in sc]
    b = capi.Batch(2, storage)
    for d in mixed:
        b.add(d)
    for d in alone:
        d.step(20)
    b.step(20)
    for a, m in zip(alone, mixed):
        same(a, m, storage)
        assert np.array_equal(a.body_wrench(), m.body_wrench())


def test_batch_3d_sphere_and_periodic():
    note("case batch3d")
    nx, ny, nz = 40, 32, 24
    m_tg = wl.taylor_green_ring(nx, ny, nz, amp=0.03, ring_r=6.0, seed=3)
    sph_v, sph_f = wl.icosphere([20.0, 16.0, 12.0], 5.0, subdiv=2)
    vel = np.zeros_like(sph_v)
    vel[:, 0] = 0.02

    def build(periodic):
        d = capi.Domain(3, nx, ny, nz, viscosity=0.02, periodic=periodic)
        d.set_moments(m_tg)
        if not periodic:
            d.set_boundary(sph_v, vel, sph_f)
        return d

    alone = [build(True), build(False)]
    mixed_p, mixed_c = build(True), build(False)
    bp = capi.Batch(3, "f32", periodic=True)
    bc = capi.Batch(3, "f32", periodic=False)
    bp.add(mixed_p)
    bc.add(mixed_c)
    for d in alone:
        d.step(12)
    bp.step(12)
    bc.step(12)
    same(alone[0], mixed_p, "3D periodic")
    same(alone[1], mixed_c, "3D sphere")
    assert np.array_equal(alone[1].body_wrench(), mixed_c.body_wrench())
    assert np.abs(alone[1].body_wrench()).max() > 0


@pytest.mark.parametrize("storage", ["f16", "bf16", "f16u"])
def test_batch_3d_16bit_storage(storage):
    # two 3D members of different sizes, each with a moving sphere, in the
    # served form's 16-bit storage: each equals its lone steps
    note(f"case batch3d storage {storage}")

    def build(nx, ny, nz, c, r, seed):
        d = capi.Domain(3, nx, ny, nz, viscosity=0.02, storage=storage)
        d.set_moments(wl.taylor_green_ring(nx, ny, nz, amp=0.03, ring_r=5.0, seed=seed))
        v, f = wl.icosphere(c, r, subdiv=2)
        vel = np.zeros_like(v)
        vel[:, 1] = 0.015
        d.set_boundary(v, vel, f)
        return d

    specs = [(40, 32, 24, [20.0, 16.0, 12.0], 5.0, 3), (33, 28, 30, [15.5, 13.2, 16.1], 4.0, 4)]
    alone = [build(*sp) for sp in specs]
    mixed = [build(*sp) for sp in specs]
    b = capi.Batch(3, storage)
    for d in mixed:
        b.add(d)
    for d in alone:
        d.step(12)
    b.step(12)
    for i, (a, m) in enumerate(zip(alone, mixed)):
        same(a, m, f"3D {storage} member {i}")
        assert np.array_equal(a.body_wrench(), m.body_wrench())
        assert np.abs(a.body_wrench()).max() > 0


def test_member_destroyed_mid_run():
    # ff_domain_destroy on a member detaches it first: the batch's graphs and
    # tables are rebuilt without it and the others go on as if stepped alone
    sc = scenes(61, n=3)
    note("case member destroyed mid-run")
    alone = [make(s) for s in sc]
    mixed = [make(s) for s in sc]
    b = capi.Batch(2, "f32")
    for d in mixed:
        b.add(d)
    b.step(9)  # 8-step graph replay plus a direct step
    for d in alone:
        d.step(9)
    mixed[1].close()
    assert len(b) == 2
    b.step(9)
    for d in (alone[0], alone[2]):
        d.step(9)
    b.sync()
    for i in (0, 2):
        same(alone[i], mixed[i], f"member {i}")
        assert np.array_equal(alone[i].body_wrench(), mixed[i].body_wrench())


def test_batch_destroyed_members_step_alone():
    # ff_batch_destroy returns every member to its own stream and step state
    sc = scenes(62, n=2)
    note("case batch destroyed, members continue alone")
    alone = [make(s) for s in sc]
    mixed = [make(s) for s in sc]
    b = capi.Batch(2, "f32")
    for d in mixed:
        b.add(d)
    b.step(11)
    b.close()
    for d in mixed:
        d.step(7)
    for d in alone:
        d.step(18)
    for i, (a, m) in enumerate(zip(alone, mixed)):
        same(a, m, f"member {i}")
        assert np.array_equal(a.body_wrench(), m.body_wrench())
        assert np.array_equal(a.wrench_history(18), m.wrench_history(18))


@pytest.mark.parametrize("batched", [False, True])
def test_link_table_overflow_is_reported(batched):
    """A boundary moved far beyond the shape its link table was sized for
    fills the table: the step reports FF_ERR_CAPACITY (alone through ff_sync,
    batched through ff_batch_sync) instead of dropping links silently."""
    nx, ny, n = 2000, 64, 20
    e = np.stack([np.arange(n), np.arange(1, n + 1)], 1)
    small = np.stack([600.0 + 0.04 * np.arange(n + 1), np.full(n + 1, 32.3)], 1)  # inside one cell
    big = np.stack([np.linspace(10.0, 1990.0, n + 1), np.full(n + 1, 32.3)], 1)  # ~4000 link cells
    note(f"case overflow batched={batched}")
    d = capi.Domain(2, nx, ny, viscosity=0.01)
    d.init_state([0.0] * 4)
    d.set_boundary(small, np.zeros_like(small), e)
    b = None
    if batched:
        b = capi.Batch(2, "f32", stage_fp16=True)
        b.add(d)

    def run(k):
        if batched:
            b.step(k)
            b.sync()
        else:
            d.step(k)
            d.sync()

    run(2)  # the shape it was sized for fits
    d.update_boundary(big, np.zeros_like(big))
    with pytest.raises(capi.FreeFlowError, match="error 4:"):
        run(1)
    if b is not None:
        b.close()
    d.close()


def test_boundary_replaced_equals_set_once():
    """ff_set_boundary on a domain that already has a boundary, a wave and a
    wrench history frees and rebuilds all of them: stepping it afterwards
    gives the same moments and wrench history as a domain set up once."""
    sc_other, sc = scenes(31, n=2)
    note(f"case boundary replaced {sc.nx}x{sc.ny}")
    d_once = capi.Domain(**sc.domain_args(), storage="f32", stage_fp16=True)
    sc.setup(d_once, history=16)
    d_twice = capi.Domain(**sc.domain_args(), storage="f32", stage_fp16=True)
    d_twice.set_flags(sc.flags)
    d_twice.init_state([0.0] * 4)
    # a first boundary (another swimmer, scaled into this grid) with wave and history
    v1, _ = sc_other.swimmer.state(0.0)
    v1 = v1 * (min(sc.nx, sc.ny) / max(sc_other.nx, sc_other.ny))
    d_twice.set_boundary(v1, np.zeros_like(v1), sc_other.swimmer.elems)
    d_twice.set_wave(sc_other.swimmer.rest, sc_other.swimmer.s, sc_other.swimmer.params[None, :])
    d_twice.set_wrench_history(8)
    # replaced before any step: the same setup as d_once
    v0, vel0 = sc.swimmer.state(0.0)
    d_twice.set_boundary(v0, vel0, sc.swimmer.elems)
    d_twice.set_wave(sc.swimmer.rest, sc.swimmer.s, sc.swimmer.params[None, :])
    d_twice.set_wrench_history(16)
    d_once.step(12)
    d_twice.step(12)
    same(d_once, d_twice, "replaced boundary")
    assert np.array_equal(d_once.wrench_history(12), d_twice.wrench_history(12))
    assert np.array_equal(d_once.vertex_forces(), d_twice.vertex_forces())
