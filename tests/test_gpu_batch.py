"""Batched graph-captured runs and the launch autotuner, through the C ABI.

A batch captures the auto-dt step of every run into one HIP graph with the
accept/reject decision on the device; it must advance each run exactly as the
run's own loop would (bit for bit: the same kernels in the same order on the
same data), across odd and even step counts, and refuse runs stepped outside
it.  The autotuner only changes launch shapes, so a tuned run must reproduce
the untuned run bit for bit, and tuning must leave the live state untouched."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NAMES = ("rho", "mx", "my", "energy", "bx_face", "by_face", "bz_cell")


def descs():
    return [{"run": 0, "n": 64, "kind": "orszag_tang", "noise": 1e-3, "noise_seed": 0},
            {"run": 1, "n": 48, "kind": "rotor", "noise": 1e-3, "noise_seed": 1},
            {"run": 2, "n": 96, "kind": "orszag_tang", "noise": 1e-3, "noise_seed": 2}]


def network():
    from quasar_amd import learned
    p, t, _ = learned.load()
    return p, t


@pytest.mark.parametrize("arithmetic,indicator", [("fp64", "none"), ("fp32", "mlp")])
def test_batch_advances_every_run_like_its_own_loop(arithmetic, indicator):
    from quasar_amd import batch
    net = network() if indicator == "mlp" else None
    ds = descs()
    batched, _ = batch.build_sweep(ds, arithmetic, indicator, net)
    alone, _ = batch.build_sweep(ds, arithmetic, indicator, net)
    b = batch.Batch(batched)
    b.run(3)
    b.run(4)  # odd then even: the register parity carries over
    for k, s in enumerate(alone):
        done, t = s.run(7)
        assert done == 7
        st = b.status(k)
        assert st == {"accepted": 7, "rejected": 0, "fatal": False, "sim_time": t}, (k, st, t)
        for nm in NAMES:
            np.testing.assert_array_equal(batched[k].state_component_to_host(nm), s.state_component_to_host(nm))


def test_batch_refuses_a_run_stepped_outside_it():
    from quasar_amd import batch
    solvers, _ = batch.build_sweep(descs()[:2], "fp64", "none")
    b = batch.Batch(solvers)
    b.run(2)
    solvers[0].run(1)
    with pytest.raises(ValueError):
        b.run(1)


def test_autotune_restores_the_state_and_keeps_results():
    from quasar_amd import batch
    d = [{"run": 0, "n": 128, "kind": "orszag_tang", "noise": 1e-3, "noise_seed": 5}]
    (tuned,), _ = batch.build_sweep(d, "fp32", "mlp", network())
    (plain,), _ = batch.build_sweep(d, "fp32", "mlp", network())
    before = [tuned.state_component_to_host(nm) for nm in NAMES]
    choice = tuned.autotune(2)
    for k, nm in enumerate(NAMES):
        np.testing.assert_array_equal(tuned.state_component_to_host(nm), before[k])
    for key in ("flux", "corner", "update"):
        bx, by = choice[f"{key}_block_x"], choice[f"{key}_block_y"]
        assert 64 <= bx * by <= 256 and (bx * by) % 64 == 0
    assert choice["list_blocks"] in (256, 512, 1024, 2048, 4096)
    assert tuned.tuning() == choice
    assert tuned.run(5)[0] == 5 and plain.run(5)[0] == 5
    for nm in NAMES:
        np.testing.assert_array_equal(tuned.state_component_to_host(nm), plain.state_component_to_host(nm))


def test_tuning_rejects_invalid_shapes():
    from quasar_amd import batch
    (s,), _ = batch.build_sweep(descs()[:1], "fp64", "none")
    for key, value in (("flux_block_x", 48), ("corner_block_y", 0), ("list_blocks", 0)):
        bad = dict(s.tuning())
        bad[key] = value
        if key == "flux_block_x":
            bad["flux_block_y"] = 1  # 48 threads: not a whole number of waves
        with pytest.raises(ValueError):
            s.set_tuning(bad)
