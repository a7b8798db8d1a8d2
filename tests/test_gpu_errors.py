"""Error returns of the C ABI on a live device: a rejected call changes
nothing, so the domain's results equal those of a domain that never saw it.
Every rejected argument here is caught on the host before any launch."""
import sys

import numpy as np
import pytest

from freeflow_amd import capi, workloads as wl

pytestmark = pytest.mark.gpu


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def domain(nx=96, ny=48):
    d = capi.Domain(2, nx, ny, viscosity=0.01)
    d.set_flags(wl.outlet_box_flags(nx, ny))
    d.set_moments(wl.equilibrium_moments(np.tile([0.02, 0.0], (nx * ny, 1))))
    return d


def test_rejected_calls_leave_the_domain_unchanged():
    note("case rejected boundary, wave and step arguments on a 96x48 domain")
    v, e = wl.cylinder(40, 40.0, 24.0, 8.0)
    vel = np.zeros_like(v)
    vel[:, 1] = 0.01
    ref, d = domain(), domain()
    # calls that need a boundary (FF_ERR_STATE = 5)
    for call in (lambda: d.update_boundary(v, vel), lambda: d.body_wrench(), lambda: d.set_wave_params(np.zeros(4))):
        with pytest.raises(capi.FreeFlowError, match="error 5:"):
            call()
    for x in (ref, d):
        x.set_boundary(v, vel, e)
    bad_e = e.copy()
    bad_e[3, 1] = 40
    nan_v = v.copy()
    nan_v[5, 0] = np.nan
    far_v = v.copy()
    far_v[5, 0] = 3e9
    many = np.zeros((65537, 2), np.uint32)
    cases = [((v, vel, bad_e), {}, "vertex out of range"),
             ((nan_v, vel, e), {}, "not finite or beyond"),
             ((far_v, vel, e), {}, "not finite or beyond"),
             ((v, vel, many), {}, "65536"),
             ((v, vel, e), dict(elem_body=np.full(40, 2), n_bodies=2), "elem_body")]
    for args, kw, msg in cases:
        with pytest.raises(capi.FreeFlowError, match=msg):
            d.set_boundary(*args, **kw)
    with pytest.raises(capi.FreeFlowError, match="error 5:"):
        d.set_wave_params(np.zeros(4))  # no wave set
    with pytest.raises(capi.FreeFlowError, match="error 1:"):
        d.step(-1)
    for x in (ref, d):
        x.step(30)
    assert np.array_equal(ref.moments(0), d.moments(0))
    assert np.array_equal(ref.body_wrench(), d.body_wrench())
    assert np.array_equal(ref.vertex_forces(), d.vertex_forces())
    assert np.abs(d.vertex_forces()).max() > 0


def test_rejected_batch_calls_leave_the_batch_unchanged():
    note("case rejected batch membership calls, two 96x48 domains with a cylinder")
    v, e = wl.cylinder(40, 40.0, 24.0, 8.0)
    vel = np.zeros_like(v)
    vel[:, 0] = 0.01
    alone, member = domain(), domain()
    for x in (alone, member):
        x.set_boundary(v, vel, e)
    b = capi.Batch(2, "f32")
    b.add(member)
    with pytest.raises(capi.FreeFlowError, match="error 5:"):
        b.add(member)  # already a member
    with pytest.raises(capi.FreeFlowError, match="differs from the batch"):
        b.add(capi.Domain(2, 32, 32, storage="f16"))
    with pytest.raises(capi.FreeFlowError, match="differs from the batch"):
        b.add(capi.Domain(3, 16, 16, 16))
    other = domain()
    with pytest.raises(capi.FreeFlowError, match="not in this batch"):
        b.remove(other)
    with pytest.raises(capi.FreeFlowError, match="error 5:"):
        member.step(1)  # a member steps with its batch
    with pytest.raises(capi.FreeFlowError, match="error 1:"):
        b.step(-1)
    with pytest.raises(capi.FreeFlowError, match="tile index"):
        b.set_tile(99)
    assert len(b) == 1
    alone.step(12)
    b.step(12)
    b.sync()
    assert np.array_equal(alone.moments(0), member.moments(0))
    assert np.array_equal(alone.body_wrench(), member.body_wrench())


@pytest.mark.parametrize("dim", [2, 3])
def test_empty_boundary_steps_like_no_boundary(dim):
    # zero vertices and elements: the plain stream-collide step, zero wrench
    note(f"case empty boundary dim={dim}")
    shape = (64, 32) if dim == 2 else (24, 16, 16)
    n = int(np.prod(shape))
    rng = np.random.default_rng(4)
    u = rng.uniform(-0.02, 0.02, (n, dim))
    plain, empty = (capi.Domain(dim, *shape, viscosity=0.01, periodic=True) for _ in range(2))
    for x in (plain, empty):
        x.set_moments(wl.equilibrium_moments(u))
    empty.set_boundary(np.zeros((0, dim), np.float32), np.zeros((0, dim), np.float32), np.zeros((0, dim), np.uint32))
    empty.set_wrench_history(4)
    for x in (plain, empty):
        x.step(6)
    assert np.array_equal(plain.moments(0), empty.moments(0))
    assert not np.any(empty.body_wrench())
    assert empty.vertex_forces().shape == (0, dim)
    assert len(empty.wrench_history(4)) == 0
