"""Restart through a context: the J2 state block after the reference's
records holds the committed state exactly as ufc_ph_j2_get_state returns it
(13 doubles per Gauss point, UMAT statev order after the stress), and a J2 job
stopped after 5 of 10 increments and resumed from its checkpoint in a fresh
context ends bit-identical to the job run straight through."""
import struct

import numpy as np
import pytest

import ufc_hip

pytestmark = pytest.mark.gpu


def _j2(n):
    from ufc_hip import problems as P
    pr = P.j2_cube(n, increments=10)
    target = np.zeros(pr.bc_eq.size)
    target[pr.extra["n_clamp"]:] = pr.extra["strain"]
    return pr, target, pr.extra["props"]


def _records(b):
    out, p = [], 0
    while p < len(b):
        (n,) = struct.unpack_from("<i", b, p)
        out.append(b[p + 4:p + 4 + n])
        p += 8 + n
    return out


@pytest.mark.parametrize("precision", [0, 2])
def test_j2_job_resumed_from_checkpoint_is_bit_identical(tmp_path, precision):
    from ufc_hip.nonlinear import j2_static
    pr, target, props = _j2(6)
    ckpt = str(tmp_path / "j2.rst")
    print(f"[gpu-test] j2 6^3 restart precision {precision}", flush=True)
    with ufc_hip.Context(0) as a:
        a.build_pattern(pr.coords, pr.conn)
        full = j2_static(a, pr.bc_eq, target, props, n_increments=10, precision=precision)
        state_full = a.j2_get_state()
    assert full.converged and len(full.increments) == 10
    with ufc_hip.Context(0) as b:
        b.build_pattern(pr.coords, pr.conn)
        first = j2_static(b, pr.bc_eq, target, props, n_increments=10, precision=precision,
                          restart_file=ckpt, stop_after=5)
    assert len(first.increments) == 5
    with ufc_hip.Context(0) as c:
        c.build_pattern(pr.coords, pr.conn)
        rest = j2_static(c, pr.bc_eq, target, props, n_increments=10, precision=precision,
                         resume_from=ckpt, restart_file=ckpt)
        state_rest = c.j2_get_state()
    assert rest.converged and len(rest.increments) == 5
    assert [i.load_factor for i in rest.increments] == \
        [i.load_factor for i in full.increments[5:]]
    assert np.array_equal(rest.u, full.u)
    assert np.array_equal(state_rest, state_full)
    # the last checkpoint holds the end of the job
    from ufc_hip import restart as R
    u_ck, s_ck = R.load(ckpt, n_dof=pr.n_dof)
    assert np.array_equal(u_ck, full.u) and s_ck.increment == 10
    assert s_ck.lambda_ == full.increments[-1].load_factor


def test_state_block_layout(ctx, tmp_path):
    from ufc_hip import restart as R
    from ufc_hip.nonlinear import j2_static
    pr, target, props = _j2(4)
    ctx.build_pattern(pr.coords, pr.conn)
    print("[gpu-test] j2 4^3 state block", flush=True)
    res = j2_static(ctx, pr.bc_eq, target, props, n_increments=4, precision=0)
    assert res.converged
    path = str(tmp_path / "s.rst")
    R.save(path, res.u, R.StepState(1.0, 4, 1.0, 0.0, 1, True), ctx=ctx, with_state=True)
    recs = _records(open(path, "rb").read())
    assert len(recs) == 7 + 1 + 1
    assert struct.unpack("<5i", recs[7]) == (pr.n_elems, 8, 6, 7, 1)
    block = np.frombuffer(recs[8], "<f8").reshape(pr.n_elems, 8, 13)
    state = ctx.j2_get_state().reshape(pr.n_elems, 8, 13)
    assert np.array_equal(block, state)
    assert block[..., 6].max() > 0.0  # eq. plastic strain (statev(1)) after yield
    # loading the file into a context with the state zeroed restores it
    ctx.j2_set_state(np.zeros(state.size))
    u2, s2 = R.load(path, ctx=ctx, with_state=True)
    assert np.array_equal(u2, res.u) and s2.increment == 4
    assert np.array_equal(ctx.j2_get_state().reshape(state.shape), state)


def test_load_errors_with_a_context(ctx, tmp_path):
    from ufc_hip import restart as R, UfcError
    pr, target, props = _j2(3)
    ctx.build_pattern(pr.coords, pr.conn)
    path = str(tmp_path / "nostate.rst")
    R.save(path, np.zeros(pr.n_dof), ctx=ctx)
    ctx.j2_init(*props)
    with pytest.raises(UfcError) as e:
        R.load(path, ctx=ctx, with_state=True)
    assert e.value.code == 2 and "no J2 state" in e.value.message
    other = str(tmp_path / "other.rst")
    R.save(other, np.zeros(7))  # host-only file of another size
    with pytest.raises(UfcError) as e:
        R.load(other, ctx=ctx)
    assert e.value.code == 1 and "holds 7 DOFs" in e.value.message


def test_truncated_state_keeps_committed_state(ctx, tmp_path):
    """A file that ends inside the J2 state records fails with
    UFC_STATUS_ERROR and leaves the context's committed state unchanged
    (the records are read into the trial copy first)."""
    from ufc_hip import restart as R, UfcError
    from ufc_hip.nonlinear import j2_static
    pr, target, props = _j2(4)
    ctx.build_pattern(pr.coords, pr.conn)
    print("[gpu-test] j2 4^3 truncated state", flush=True)
    res = j2_static(ctx, pr.bc_eq, target, props, n_increments=4, precision=0)
    assert res.converged
    path = str(tmp_path / "s.rst")
    R.save(path, res.u, R.StepState(1.0, 4, 1.0, 0.0, 1, True), ctx=ctx, with_state=True)
    data = open(path, "rb").read()
    cut = str(tmp_path / "cut.rst")
    open(cut, "wb").write(data[:len(data) - 4 - 8 * 13 * 8 * 5])  # inside the last record
    known = np.random.default_rng(3).normal(size=pr.n_elems * 8 * 13)
    ctx.j2_set_state(known)
    with pytest.raises(UfcError) as e:
        R.load(cut, ctx=ctx, with_state=True)
    assert e.value.code == 2
    assert np.array_equal(ctx.j2_get_state().ravel(), known)
