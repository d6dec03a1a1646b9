"""Learned form of the MHD step: training data, training loop and the stored
troubled-cell indicator.

The indicator is a 100-32-32-1 ReLU network over the 5x5 stencil features of
(rho, p, |B|, |v|) of a cell (include/quasar_mhd.h, qmhd_cell_features).  It is
trained on labels taken from the plain method itself (qmhd_cell_labels: an MP
limiter clamp, the reconstruction fallback chain, or an HLLD-vs-HLL flux gap
larger than LABEL_TOL on any face of the cell) in seeded
multi-blast runs, with the library's own BF16-MFMA forward/backward kernels
and Adam.  Device buffers are torch tensors (plumbing only: every arithmetic
step runs in the library's kernels).

Command line (needs a HIP device):
    python -m quasar_amd.learned --out quasar_amd/data/indicator_mlp.npz
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import pathlib
import sys
import time
from typing import Dict, List, Tuple

import numpy as np

from . import _lib, mhd

PARAMS = _lib.MLP_PARAMS
FEATURES = _lib.MLP_FEATURES
HIDDEN = _lib.MLP_HIDDEN
DEFAULT_WEIGHTS = pathlib.Path(__file__).resolve().parent / "data" / "indicator_mlp.npz"
LABEL_BITS = _lib.LABEL_MP_CLIP | _lib.LABEL_MUSCL | _lib.LABEL_CONSTANT | _lib.LABEL_HLL_GAP
LABEL_TOL = 1e-6

# Offsets of the parameter vector (W1[100][32], b1, W2[32][32], b2, w3, b3).
OFF_W1, OFF_B1 = 0, FEATURES * HIDDEN
OFF_W2 = OFF_B1 + HIDDEN
OFF_B2 = OFF_W2 + HIDDEN * HIDDEN
OFF_W3 = OFF_B2 + HIDDEN
OFF_B3 = OFF_W3 + HIDDEN


def rng_for(*key: int) -> np.random.Generator:
    """Counter-based generator (Philox-4x32-10) keyed by integers."""
    k = 0
    for v in key:
        k = (k * 1_000_003 + int(v)) % (1 << 64)
    return np.random.Generator(np.random.Philox(key=k))


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal weights, zero biases."""
    r = rng_for(seed, 17)
    p = np.zeros(PARAMS, dtype=np.float32)
    p[OFF_W1:OFF_B1] = r.normal(0.0, math.sqrt(2.0 / FEATURES), FEATURES * HIDDEN)
    p[OFF_W2:OFF_B2] = r.normal(0.0, math.sqrt(2.0 / HIDDEN), HIDDEN * HIDDEN)
    p[OFF_W3:OFF_B3] = r.normal(0.0, math.sqrt(1.0 / HIDDEN), HIDDEN)
    return p


def save(path, params: np.ndarray, threshold: float, meta: Dict) -> None:
    path = pathlib.Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    np.savez(path, params=np.asarray(params, dtype=np.float32), threshold=np.float64(threshold),
             meta=np.array(json.dumps(meta)))


def load(path=DEFAULT_WEIGHTS) -> Tuple[np.ndarray, float, Dict]:
    with np.load(path, allow_pickle=False) as z:
        params = z["params"].astype(np.float32)
        threshold = float(z["threshold"])
        meta = json.loads(str(z["meta"]))
    if params.shape != (PARAMS,):
        raise ValueError(f"{path}: indicator parameters must have shape ({PARAMS},)")
    return params, threshold, meta


# ---------------------------------------------------------------------------
# Training runs: seeded multi-blast problems.

def multi_blast(n: int, run: int, seed: int = 0, ng: int = 4):
    """Planes of a seeded multi-blast problem on [0,1]^2: rho = 1, v = 0, a
    uniform field B = (Bx, 0, 0) with Bx ~ U[0,1], ambient p = 0.1 and 1-8
    circular blasts (radius U[0.04, 0.12]) at p = 0.1 * U[10, 100].  Returns
    (gamma, planes in the solver's component order, description)."""
    r = rng_for(seed, run, 29)
    count = int(r.integers(1, 9))
    centres = r.uniform(0.15, 0.85, size=(count, 2))
    radii = r.uniform(0.04, 0.12, size=count)
    ratios = r.uniform(10.0, 100.0, size=count)
    bx = float(r.uniform(0.0, 1.0))
    gamma = 5.0 / 3.0 if run % 2 == 0 else 1.4
    p_amb = 0.1
    h = 1.0 / n
    c = (np.arange(-ng, n + ng) + 0.5) * h
    x, y = np.meshgrid(c, c)
    p = np.full_like(x, p_amb)
    for (cx, cy), rad, ratio in zip(centres, radii, ratios):
        inside = (x - cx) ** 2 + (y - cy) ** 2 < rad * rad
        p = np.where(inside, np.maximum(p, p_amb * ratio), p)
    rho = np.ones_like(x)
    zero = np.zeros_like(x)
    energy = p / (gamma - 1.0) + 0.5 * bx * bx
    planes = [rho, zero, zero, zero, energy, np.full_like(x, bx), zero, zero]
    desc = dict(run=run, blasts=count, bx=bx, gamma=gamma, p_ratio_max=float(ratios.max()))
    return gamma, [np.ascontiguousarray(a.ravel()) for a in planes], desc


def _solver(n: int, gamma: float, sides: str, arithmetic: str = "fp64", indicator: str = "none", ng: int = 4):
    cfg = mhd.MhdConfig()
    cfg.grid = mhd.Grid2D(n, n, 1.0, 1.0, 0.0, 0.0, ng)
    cfg.gamma = gamma
    cfg.reconstruction = "mp7"
    cfg.arithmetic = arithmetic
    cfg.indicator = indicator
    cfg.boundary.set_fluid_all(sides)
    cfg.boundary.set_field_all(sides)
    return mhd.MhdSolver2D(cfg)


def snapshot_arrays(solver: "mhd.MhdSolver2D", n: int, label_tol: float = LABEL_TOL):
    """(features float32 [n*n,100], label bits uint8 [n*n]) of the live state."""
    import torch
    feats = torch.empty((n * n, FEATURES), dtype=torch.float32, device="cuda")
    bits = torch.empty(n * n, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    solver.cell_features_into(feats.data_ptr())
    solver.cell_labels_into(label_tol, bits.data_ptr())
    return feats.cpu().numpy(), bits.cpu().numpy()


def build_dataset(runs: int, n: int = 512, snapshots=(8, 24, 48), seed: int = 0, neg_ratio: float = 3.0,
                  log=print):
    """Features and labels from `runs` multi-blast runs (plain FP64 MP7),
    sampled at the given step counts: every troubled cell and up to
    neg_ratio times as many smooth cells per snapshot."""
    xs, ys = [], []
    for run in range(runs):
        gamma, planes, desc = multi_blast(n, run, seed)
        s = _solver(n, gamma, "outflow")
        for k, name in enumerate(mhd.STATE_COMPONENTS):
            s.seed_state(name, planes[k])
        done = 0
        for target in snapshots:
            got, _ = s.run(target - done)
            done += got
            f, b = snapshot_arrays(s, n)
            lab = (b & LABEL_BITS) != 0
            pos = np.flatnonzero(lab)
            neg = np.flatnonzero(~lab)
            r = rng_for(seed, run, target, 31)
            take = min(len(neg), int(neg_ratio * max(len(pos), 1)))
            neg = r.choice(neg, size=take, replace=False) if take < len(neg) else neg
            idx = np.concatenate([pos, neg])
            xs.append(f[idx])
            ys.append(lab[idx].astype(np.float32))
            log(f"learned: run {run} ({desc['blasts']} blasts, gamma {gamma:.3f}) step {done}: "
                f"{len(pos)} troubled / {n * n} cells")
    return np.concatenate(xs), np.concatenate(ys)


# ---------------------------------------------------------------------------
# Training on the device with the library's kernels.

class Trainer:
    """Adam on weighted binary cross-entropy; minibatches of a fixed shuffled
    order per epoch (seeded)."""

    def __init__(self, x: np.ndarray, y: np.ndarray, params: np.ndarray, pos_weight: float = 2.0,
                 chunks: int = 4, seed: int = 0):
        import torch
        self.torch = torch
        self.lib = _lib.load()
        self.n = int(x.shape[0])
        self.chunks = chunks
        self.seed = seed
        npos = float(y.sum())
        nneg = float(self.n - npos)
        # Class-balanced weights (positives emphasised), normalised per sample.
        w = np.where(y > 0.5, pos_weight * 0.5 / max(npos, 1.0), 0.5 / max(nneg, 1.0)).astype(np.float32)
        self.base_w = w * self.n / (pos_weight * 0.5 + 0.5)
        dev = "cuda"
        self.x = torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).to(dev)
        self.y = torch.from_numpy(np.ascontiguousarray(y, dtype=np.float32)).to(dev)
        self.w = torch.from_numpy(self.base_w).to(dev)
        self.p = torch.from_numpy(np.ascontiguousarray(params, dtype=np.float32)).to(dev)
        self.m = torch.zeros(PARAMS, dtype=torch.float32, device=dev)
        self.v = torch.zeros(PARAMS, dtype=torch.float32, device=dev)
        self.packed = torch.zeros(_lib.MLP_PACKED, dtype=torch.int16, device=dev)
        self.grad = torch.zeros(PARAMS + 1, dtype=torch.float32, device=dev)
        self.partial = None
        self.step = 0
        torch.cuda.synchronize()

    def _ptr(self, t, offset_elems: int = 0) -> C.c_void_p:
        return C.c_void_p(t.data_ptr() + offset_elems * t.element_size())

    def epoch(self, batch: int, lr: float, epoch_index: int) -> float:
        torch = self.torch
        order = rng_for(self.seed, epoch_index, 37).permutation(self.n).astype(np.int64)
        order_dev = torch.from_numpy(order).to("cuda")
        wscale = 1.0 / batch
        wdev = torch.from_numpy(self.base_w * wscale).to("cuda")
        need = self.lib.qmhd_mlp_partial_floats(batch, self.chunks)
        if self.partial is None or self.partial.numel() < need:
            self.partial = torch.zeros(need, dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        total = 0.0
        for b0 in range(0, self.n, batch):
            nb = min(batch, self.n - b0)
            _lib.check(self.lib.qmhd_mlp_pack(self._ptr(self.p), self._ptr(self.packed), None))
            _lib.check(self.lib.qmhd_mlp_gradient(self._ptr(self.packed), self._ptr(self.p), self._ptr(self.x),
                                                  self._ptr(self.y), self._ptr(wdev), self._ptr(order_dev, b0),
                                                  nb, self.chunks, self._ptr(self.partial), self._ptr(self.grad),
                                                  None))
            self.step += 1
            _lib.check(self.lib.qmhd_mlp_adam(self._ptr(self.p), self._ptr(self.grad), self._ptr(self.m),
                                              self._ptr(self.v), self.step, lr, 0.9, 0.999, 1e-8, None))
        _lib.check(self.lib.qmhd_stream_synchronize(None))
        # Loss of the last minibatch (a progress figure).
        total = float(self.grad[PARAMS].cpu())
        return total

    def params(self) -> np.ndarray:
        _lib.check(self.lib.qmhd_stream_synchronize(None))
        return self.p.cpu().numpy().copy()


def device_logits(params: np.ndarray, x: np.ndarray) -> np.ndarray:
    """Logits of the rows of x with the library's forward kernel."""
    import torch
    lib = _lib.load()
    p = torch.from_numpy(np.ascontiguousarray(params, dtype=np.float32)).cuda()
    xd = torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).cuda()
    packed = torch.zeros(_lib.MLP_PACKED, dtype=torch.int16, device="cuda")
    out = torch.zeros(x.shape[0], dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    _lib.check(lib.qmhd_mlp_pack(C.c_void_p(p.data_ptr()), C.c_void_p(packed.data_ptr()), None))
    _lib.check(lib.qmhd_mlp_forward(C.c_void_p(packed.data_ptr()), C.c_void_p(p.data_ptr()),
                                    C.c_void_p(xd.data_ptr()), x.shape[0], C.c_void_p(out.data_ptr()), None))
    _lib.check(lib.qmhd_stream_synchronize(None))
    return out.cpu().numpy()


def choose_threshold(logits: np.ndarray, labels: np.ndarray, recall: float) -> float:
    """Largest logit threshold that keeps at least `recall` of the troubled
    samples at or above it (flagged)."""
    pos = np.sort(logits[labels > 0.5])
    if len(pos) == 0:
        return float("inf")
    k = int(math.floor((1.0 - recall) * len(pos)))
    return float(np.nextafter(np.float32(pos[k]), np.float32(-np.inf)))


def train(runs: int = 8, n: int = 512, epochs: int = 20, batch: int = 4096, lr: float = 2e-3, recall: float = 0.998,
          seed: int = 0, log=print):
    t0 = time.time()
    x, y = build_dataset(runs, n, seed=seed, log=log)
    r = rng_for(seed, 41)
    perm = r.permutation(len(y))
    nval = len(y) // 5
    val, tr = perm[:nval], perm[nval:]
    log(f"learned: dataset {len(y)} samples ({int(y.sum())} troubled), built in {time.time() - t0:.1f} s")
    trainer = Trainer(x[tr], y[tr], init_params(seed), seed=seed)
    history = []
    for e in range(epochs):
        lr_e = lr * (0.5 * (1 + math.cos(math.pi * e / epochs)))
        loss = trainer.epoch(batch, lr_e, e)
        history.append(loss)
        log(f"learned: epoch {e} lr {lr_e:.2e} last-batch loss {loss:.5f}")
    params = trainer.params()
    zv = device_logits(params, x[val])
    thr = choose_threshold(zv, y[val], recall)
    flagged = zv >= thr
    yv = y[val] > 0.5
    metrics = dict(val_samples=int(nval), val_troubled=int(yv.sum()),
                   recall=float((flagged & yv).sum() / max(yv.sum(), 1)),
                   smooth_sent_cheap=float((~flagged & ~yv).sum() / max((~yv).sum(), 1)),
                   accuracy_at_zero=float(((zv >= 0) == yv).mean()), threshold_logit=thr,
                   train_seconds=time.time() - t0, history=history)
    meta = dict(runs=runs, n=n, epochs=epochs, batch=batch, lr=lr, recall_target=recall, seed=seed,
                label_tol=LABEL_TOL, label_bits=LABEL_BITS, **metrics)
    return params, thr, meta


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="Train the troubled-cell indicator on the GPU")
    ap.add_argument("--runs", type=int, default=8)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--recall", type=float, default=0.998)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=str(DEFAULT_WEIGHTS))
    a = ap.parse_args(argv)
    import torch
    if not torch.cuda.is_available():
        print("learned: no HIP device", file=sys.stderr)
        return 1
    torch.cuda.init()
    params, thr, meta = train(a.runs, a.n, a.epochs, a.batch, a.lr, a.recall, a.seed,
                              log=lambda s: print(s, flush=True))
    save(a.out, params, thr, meta)
    print(json.dumps({k: v for k, v in meta.items() if k != "history"}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
