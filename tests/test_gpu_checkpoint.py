"""Checkpoint/resume of the ADMM state (an extension; the reference has no counterpart).

A run checkpointed after k iterations and resumed in a fresh solver (same mesh and
settings) must continue bit for bit: the same positions and the same per-iteration logs
as the uninterrupted run, across a reweighting (the reweight delay is 50 iterations) and,
for the mixed path, across the switch from FP32 to FP64 local steps.
"""
import numpy as np
import pytest

from wrapd_amd import meshgen
from wrapd_amd.solver import Device, Solver, WrapdError

pytestmark = pytest.mark.gpu


def _iterate(s, n):
    for _ in range(n):
        s.iterate()


@pytest.mark.parametrize("path", ["fp64_lbfgs", "fp32_newton"])
def test_resume_continues_bit_for_bit(tmp_path, path):
    prob = meshgen.small_beam(20, 4, 4, seed=1, pin_far_end=False)
    over = {}
    k, more = 45, 25
    if path == "fp32_newton":   # checkpoint a few iterations before the FP64 switch
        sw = -1
        for thr in (1e-4, 1e-3, 1e-2):   # the first switch threshold the run crosses
            over = dict(precision=1, prox_mode=1, fp64_switch=thr)
            p = Solver(prob, **over)
            p.initialize()
            p.setup()
            _iterate(p, k + more)
            sw = int(p.runtime()["fp64_switch_iter"])
            p.close()
            if sw > 4:
                break
        assert sw > 4, "the mixed run should switch to FP64 within the test's iterations"
        k = sw - 3
        more = max(more, 55 - k)   # still past the reweight delay
    ck = tmp_path / "run.ckpt"

    a = Solver(prob, **over)
    a.initialize()
    a.setup()
    _iterate(a, k)
    a.checkpoint(ck)
    _iterate(a, more)
    a.finish()

    b = Solver(prob, **over)
    b.initialize()
    b.setup()
    b.resume(ck)
    _iterate(b, more)
    b.finish()

    assert np.array_equal(a.x(), b.x())
    for name in ("objectives", "reweighted", "inner_iters"):
        la, lb = a.log(name), b.log(name)
        assert len(la) == len(lb) == k + more + 1 or name == "inner_iters"
        assert np.array_equal(la, lb), name
    assert max(a.log("reweighted")[k + 1:]) > 0   # the resumed stretch reweights
    ra, rb = a.runtime(), b.runtime()
    assert ra["fp64_switch_iter"] == rb["fp64_switch_iter"]
    if path == "fp32_newton":
        assert ra["fp64_switch_iter"] > k   # the switch happens after the checkpoint
    a.close()
    b.close()


def test_resume_rejects_another_mesh(tmp_path):
    ck = tmp_path / "run.ckpt"
    a = Solver(meshgen.small_beam(8, 2, 2, seed=3))
    a.initialize()
    a.setup()
    _iterate(a, 3)
    a.checkpoint(ck)
    b = Solver(meshgen.small_beam(9, 2, 2, seed=3))
    b.initialize()
    b.setup()
    with pytest.raises(Exception, match="another mesh"):
        b.resume(ck)
    a.close()
    b.close()


def test_device_state_round_trip():
    prob = meshgen.small_beam(10, 3, 2, seed=12)
    d = Device()
    d.upload_mesh(prob)
    mask = np.ones(prob.n_verts, bool)
    mask[prob.pins] = False
    d.set_x_fix(prob.init[np.sort(prob.pins)])
    d.set_x_free(prob.init[mask])
    d.init_duals()
    blob = d.state_save()
    u0, z0 = d.elem("U").copy(), d.elem("Z").copy()
    d.set_x_free(prob.rest[mask])
    d.init_duals()                 # changes the duals
    assert not np.array_equal(d.elem("Z"), z0)
    d.state_load(blob)
    assert np.array_equal(d.elem("U"), u0) and np.array_equal(d.elem("Z"), z0)
    with pytest.raises(Exception):
        d.state_load(blob[:-8])    # wrong size


def test_state_load_rejects_corrupt_bookkeeping():
    """wrapd_state_load checks the L-BFGS bookkeeping it will index device buffers with: the
    slot map must be a permutation of the 8 slots, the pair count within the memory."""
    prob = meshgen.small_beam(8, 2, 2, seed=3)
    s = Solver(prob)
    s.initialize()
    s.setup()
    _iterate(s, 3)
    d = s.device()
    blob = bytearray(d.state_save())
    # header: 8-byte magic, 9 ints (n_e, n_ep, n_f, n_fix, nhist, rot_aware, m_k, last_flips,
    # grad_old_valid), then the 8-int slot map
    bad = bytearray(blob)
    bad[44:48] = (9).to_bytes(4, "little", signed=True)      # slot 9 of 8
    with pytest.raises(WrapdError) as e:
        d.state_load(bytes(bad))
    assert e.value.status == 1
    bad = bytearray(blob)
    bad[32:36] = (100).to_bytes(4, "little", signed=True)    # m_k beyond the memory
    with pytest.raises(WrapdError) as e:
        d.state_load(bytes(bad))
    assert e.value.status == 1
    d.state_load(bytes(blob))                                 # the genuine blob still loads
