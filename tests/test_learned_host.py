"""Host-side logic of the learned form and the batched sweep (no GPU): the
seeded multi-blast training problems, the sweep's run list, the threshold
rule, parameter initialisation and the stored network's file format."""
import json

import numpy as np
import pytest

from quasar_amd import batch, learned


def test_multi_blast_is_seeded_and_matches_its_description():
    g1, p1, d1 = learned.multi_blast(64, 3, seed=5)
    g2, p2, d2 = learned.multi_blast(64, 3, seed=5)
    assert g1 == g2 and d1 == d2 and all(np.array_equal(a, b) for a, b in zip(p1, p2))
    _, p3, _ = learned.multi_blast(64, 4, seed=5)
    assert not np.array_equal(p1[4], p3[4])
    assert 1 <= d1["blasts"] <= 8 and 0.0 <= d1["bx"] <= 1.0 and d1["gamma"] in (5 / 3, 1.4)
    n, ng = 64, 4
    assert all(a.size == (n + 2 * ng) ** 2 for a in p1)
    rho, mx, my, mz, e, bxf, byf, bz = p1
    assert np.all(rho == 1.0) and not mx.any() and not my.any() and not mz.any() and not byf.any()
    assert np.all(bxf == d1["bx"])  # uniform field: discretely divergence free
    p = (e - 0.5 * d1["bx"] ** 2) * (d1["gamma"] - 1.0)
    assert np.isclose(p.min(), 0.1) and p.max() <= 0.1 * d1["p_ratio_max"] * (1 + 1e-12)
    assert p.max() >= 1.0  # at least one blast at ratio >= 10


def test_sweep_runs_are_seeded_and_alternate():
    a = batch.sweep_runs(64, seed=0)
    b = batch.sweep_runs(64, seed=0)
    assert a == b
    assert [d["kind"] for d in a[:4]] == ["orszag_tang", "rotor", "orszag_tang", "rotor"]
    assert {d["n"] for d in a} <= set(batch.SWEEP_SIZES)
    assert len({d["n"] for d in a}) == len(batch.SWEEP_SIZES)  # 64 draws hit every size
    assert all(d["noise"] == 1e-3 and d["noise_seed"] == d["run"] for d in a)
    assert batch.sweep_runs(64, seed=1) != a


def test_threshold_keeps_the_requested_recall():
    rng = np.random.Generator(np.random.Philox(key=3))
    pos = rng.normal(2.0, 1.0, 1000)
    neg = rng.normal(-2.0, 1.0, 3000)
    z = np.concatenate([pos, neg]).astype(np.float32)
    y = np.concatenate([np.ones(1000), np.zeros(3000)]).astype(np.float32)
    for recall in (0.9, 0.99, 0.998):
        t = learned.choose_threshold(z, y, recall)
        assert (z[y > 0.5] >= t).mean() >= recall
    assert learned.choose_threshold(z[y < 0.5], y[y < 0.5], 0.99) == float("inf")


def test_init_params_shapes_and_scales():
    p = learned.init_params(7)
    assert p.shape == (learned.PARAMS,) and p.dtype == np.float32
    w1 = p[learned.OFF_W1:learned.OFF_B1]
    assert abs(w1.std() - np.sqrt(2.0 / 100)) < 0.01
    assert not p[learned.OFF_B1:learned.OFF_W2].any() and not p[learned.OFF_B2:learned.OFF_W3].any()
    assert np.array_equal(p, learned.init_params(7))


def test_network_file_roundtrip(tmp_path):
    p = learned.init_params(1)
    path = tmp_path / "net.npz"
    learned.save(path, p, -1.25, {"runs": 2, "note": "x"})
    q, t, meta = learned.load(path)
    assert np.array_equal(p, q) and t == -1.25 and meta == {"runs": 2, "note": "x"}
    with np.load(path, allow_pickle=False) as z:
        assert json.loads(str(z["meta"]))["runs"] == 2


def test_committed_network_is_well_formed():
    p, t, meta = learned.load()
    assert p.shape == (learned.PARAMS,) and np.isfinite(p).all() and np.isfinite(t)
    assert meta["recall"] >= meta["recall_target"] - 1e-3
    assert meta["label_tol"] == learned.LABEL_TOL and meta["label_bits"] == learned.LABEL_BITS


def test_wrong_parameter_count_is_rejected(tmp_path):
    path = tmp_path / "bad.npz"
    np.savez(path, params=np.zeros(10, np.float32), threshold=np.float64(0.0), meta=np.array("{}"))
    with pytest.raises(ValueError):
        learned.load(path)
