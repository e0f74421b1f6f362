"""GPU: exact restart of the fractional step.

The state of an AB2 run between two steps is the velocity, the pressure (with
its border unknown), the step counter and the explicit source of the last step
(S_old, fractional_step_1.cpp:120-134).  Handing all of it to a fresh context
(mp_fields_set with MP_SU/MP_SV) must continue the run bit for bit; the
reference's tecplot restart carries no S_old and starts AB2 anew, which this
test shows is not the same run.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _print(*a):
    print(*a, flush=True)


def _context(M, cloud, dt, re, precision):
    ctx = M.Context(0)
    ctx.upload(cloud)
    ctx.weights(3, 3, download=False)
    u_dir, p_dir, u0 = M.lid_conditions(cloud)
    ctx.fs_create(1.0, 1.0 / re, dt, 2, 1e-10, 20000, u_dir, u_dir, None, p_dir,
                  precision=precision)
    return ctx, u0


@pytest.mark.parametrize("precision", [0, 1])
def test_restart_with_source_history_is_exact(precision):
    import memphys_amd as M
    cloud = M.box_cloud(dim=2, n=36, jitter=0.25, seed=31, k_interior=30)
    dt, re, n1, n2 = 2e-3, 100.0, 3, 3
    fields = (M.MP_U, M.MP_V, M.MP_P)
    # reference run: n1 + n2 steps in one context
    a, u0 = _context(M, cloud, dt, re, precision)
    a.set_field(M.MP_U, u0)
    a.set_field(M.MP_V, np.zeros(cloud.nv))
    a.set_field(M.MP_P, np.zeros(a.n_sys))
    a.step(n1, 0)
    state = {w: a.get_field(w) for w in fields + (M.MP_SU, M.MP_SV)}
    a.step(n2, n1)
    straight = [a.get_field(w) for w in fields]
    a.close()
    # restart: a fresh context given the state after n1 steps
    b, _ = _context(M, cloud, dt, re, precision)
    for w, val in state.items():
        b.set_field(w, val)
    b.step(n2, n1)
    restarted = [b.get_field(w) for w in fields]
    b.close()
    same = [np.array_equal(x, y) for x, y in zip(straight, restarted)]
    _print(f"precision {precision}: restart with S_old identical {same}")
    assert all(same)
    # without the source history (as a tecplot restart): a different run
    c, _ = _context(M, cloud, dt, re, precision)
    for w in fields:
        c.set_field(w, state[w])
    c.step(n2, n1)
    dropped = [c.get_field(w) for w in fields]
    c.close()
    diff = max(np.max(np.abs(x - y)) / np.max(np.abs(x)) for x, y in zip(straight, dropped))
    _print(f"precision {precision}: restart without S_old differs by {diff:.3e}")
    assert diff > 1e-8


def test_history_fields_checked():
    import memphys_amd as M
    cloud = M.box_cloud(dim=2, n=12, jitter=0.25, seed=3, k_interior=20)
    ctx, _ = _context(M, cloud, 1e-3, 100.0, 0)
    with pytest.raises(M.MemphysError):
        ctx.get_field(M.MP_SW)  # no third component in 2D
    with pytest.raises(M.MemphysError):
        ctx.get_field(7)
    ctx.close()
