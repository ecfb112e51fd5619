"""The learned seam: decision columns, classifier training on the GPU, classifier files.

The reference corrects an anchor through a classifier seam when asked to
(correctAnchorClf, include/corrector.hpp:1448-1472): HQ anchors take the consensus; in the
other anchors, each anchor column whose base differs from the consensus is set to the
consensus only if a forest (include/forest.hpp) says so. Here the classifier is an MLP
16 -> 64 -> 64 -> 1 whose logit decides when its confidence is at least 0.9 and whose
undecided columns fall back to the classic rule (getCorrectedAnchor, src/msa.cpp:379-453).

This module
- collects decision columns and their 16 features from the device
  (``care_hip_extract_features``) and labels them with the simulator's truth
  (1 = the consensus base is the true base);
- trains the MLP on the GPU with the library's forward/backward kernels and Adam
  (``care_mlp_*``), on standardized features, and folds the standardization into the
  first layer so the device sees raw features;
- reads and writes classifier files: b"CAREMLP1", u32 version 1, u32 parameter count,
  u32 length of a JSON description, the JSON, then the float32 parameters (W1[64][16],
  b1[64], W2[64][64], b2[64], W3[64], b3), little endian.
"""
from __future__ import annotations

import ctypes
import json
import os
import struct

import numpy as np

from . import engine

MAGIC = b"CAREMLP1"
N_PARAMS = engine.MLP_PARAMS
N_FEATURES = 16
HIDDEN = 64
OFF_W1, OFF_B1 = 0, HIDDEN * N_FEATURES
OFF_W2 = OFF_B1 + HIDDEN
OFF_B2 = OFF_W2 + HIDDEN * HIDDEN
OFF_W3 = OFF_B2 + HIDDEN
OFF_B3 = OFF_W3 + HIDDEN
DEFAULT_CLASSIFIER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "seam_classifier.bin")

FEATURE_NAMES = [
    "orig_coverage/coverage", "orig_weight/weight_sum", "orig_weight/orig_coverage", "cons_count/coverage",
    "support", "cons_weight/cons_count", "weight_sum/coverage", "coverage/est_coverage",
    "weight_sum/est_coverage", "avg_support", "min_support", "max_coverage/est_coverage",
    "min_coverage/est_coverage", "anchor_quality_weight", "position/length", "min_neighbour_support",
]


class Mlp:
    """Device-resident classifier parameters, Adam state and training rows (care_mlp_*)."""

    def __init__(self, device: int = 0):
        self.lib = engine.load_library()
        h = ctypes.c_void_p()
        rc = self.lib.care_mlp_create(device, ctypes.byref(h))
        if rc != engine.CARE_OK:
            raise engine.CareError(f"care_mlp_create failed with code {rc}")
        self.h = h

    def _check(self, rc, what):
        if rc != engine.CARE_OK:
            msg = self.lib.care_mlp_last_error(self.h).decode(errors="replace")
            raise engine.CareError(f"{what} failed ({rc}): {msg}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.care_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params):
        p = np.ascontiguousarray(params, np.float32)
        self._check(self.lib.care_mlp_set_params(self.h, p.ctypes.data, p.size), "care_mlp_set_params")

    def get_params(self) -> np.ndarray:
        p = np.zeros(N_PARAMS, np.float32)
        self._check(self.lib.care_mlp_get_params(self.h, p.ctypes.data, p.size), "care_mlp_get_params")
        return p

    def forward(self, X) -> np.ndarray:
        X = np.ascontiguousarray(X, np.float32)
        z = np.zeros(len(X), np.float32)
        self._check(self.lib.care_mlp_forward(self.h, X.ctypes.data, len(X), z.ctypes.data), "care_mlp_forward")
        return z

    def loss_grad(self, X, y):
        X = np.ascontiguousarray(X, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        loss = ctypes.c_float(0.0)
        g = np.zeros(N_PARAMS, np.float32)
        self._check(self.lib.care_mlp_loss_grad(self.h, X.ctypes.data, y.ctypes.data, len(X), ctypes.byref(loss),
                                                g.ctypes.data), "care_mlp_loss_grad")
        return float(loss.value), g

    def set_data(self, X, y):
        X = np.ascontiguousarray(X, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        self._check(self.lib.care_mlp_set_data(self.h, X.ctypes.data, y.ctypes.data, len(X)), "care_mlp_set_data")

    def reset_optimizer(self):
        self._check(self.lib.care_mlp_reset_optimizer(self.h), "care_mlp_reset_optimizer")

    def train(self, epochs: int, batch: int = 65536, lr: float = 1e-3, seed: int = 0) -> np.ndarray:
        losses = np.zeros(max(epochs, 1), np.float32)
        self._check(self.lib.care_mlp_train(self.h, epochs, batch, lr, seed, losses.ctypes.data), "care_mlp_train")
        return losses[:epochs]


def init_params(seed: int = 0) -> np.ndarray:
    """He-uniform weights for the ReLU layers, Glorot-uniform for the output, zero biases."""
    rng = np.random.default_rng(seed)
    p = np.zeros(N_PARAMS, np.float32)
    p[OFF_W1:OFF_B1] = rng.uniform(-1, 1, HIDDEN * N_FEATURES) * np.sqrt(6.0 / N_FEATURES)
    p[OFF_W2:OFF_B2] = rng.uniform(-1, 1, HIDDEN * HIDDEN) * np.sqrt(6.0 / HIDDEN)
    p[OFF_W3:OFF_B3] = rng.uniform(-1, 1, HIDDEN) * np.sqrt(6.0 / (HIDDEN + 1))
    return p


def labels_from_truth(meta: np.ndarray, truth: np.ndarray) -> np.ndarray:
    """1 where the consensus base of the decision column is the true base."""
    cons = np.frombuffer(b"ACGT", np.uint8)[(meta[:, 2] >> 2) & 3]
    true = truth[meta[:, 0].astype(np.int64), meta[:, 1].astype(np.int64)]
    return (cons == true).astype(np.float32)


def collect(s, rs, device: int = 0, **opts):
    """Decision columns of every anchor of a simulated storage: (X [m,16], y [m], meta [m,4])."""
    opts.setdefault("paired", int(rs.paired))
    opts.setdefault("use_quality_scores", int(rs.quals is not None))
    eng = engine.CareEngine(device, **opts)
    try:
        eng.set_reads(rs)
        eng.build_tables()
        X, meta = eng.extract_features()
    finally:
        eng.close()
    return X, labels_from_truth(meta, s.truth), meta


def fold_standardization(params: np.ndarray, mean: np.ndarray, std: np.ndarray) -> np.ndarray:
    """Parameters for raw features x from parameters trained on (x - mean) / std."""
    p = np.array(params, np.float64)
    W1 = p[OFF_W1:OFF_B1].reshape(HIDDEN, N_FEATURES)
    b1 = p[OFF_B1:OFF_W2]
    W1f = W1 / std[None, :]
    b1f = b1 - W1f @ mean
    p[OFF_W1:OFF_B1] = W1f.ravel()
    p[OFF_B1:OFF_W2] = b1f
    return p.astype(np.float32)


def train_classifier(X, y, epochs: int = 60, batch: int = 65536, lr: float = 1e-3, seed: int = 0,
                     device: int = 0):
    """Trains on the GPU; returns (parameters for raw features, per-epoch losses, info)."""
    X = np.asarray(X, np.float32)
    mean = X.mean(0).astype(np.float64)
    std = X.std(0).astype(np.float64)
    std[std < 1e-6] = 1.0
    Xs = ((X - mean) / std).astype(np.float32)
    mlp = Mlp(device)
    try:
        mlp.set_params(init_params(seed))
        mlp.set_data(Xs, y)
        losses = mlp.train(epochs, batch, lr, seed)
        trained = mlp.get_params()
    finally:
        mlp.close()
    info = dict(rows=int(len(X)), positives=float(np.mean(y)) if len(y) else 0.0, epochs=epochs, batch=batch,
                lr=lr, seed=seed, first_loss=float(losses[0]) if len(losses) else None,
                final_loss=float(losses[-1]) if len(losses) else None,
                feature_mean=mean.tolist(), feature_std=std.tolist())
    return fold_standardization(trained, mean, std), losses, info


def save_classifier(path: str, params, info: dict | None = None) -> None:
    p = np.ascontiguousarray(params, np.float32)
    if p.size != N_PARAMS:
        raise ValueError(f"expected {N_PARAMS} parameters, got {p.size}")
    meta = json.dumps(info or {}, sort_keys=True).encode()
    with open(path, "wb") as fh:
        fh.write(MAGIC + struct.pack("<III", 1, N_PARAMS, len(meta)) + meta + p.astype("<f4").tobytes())


def load_classifier(path: str = DEFAULT_CLASSIFIER):
    """(parameters, info) of a classifier file."""
    with open(path, "rb") as fh:
        blob = fh.read()
    if blob[:8] != MAGIC:
        raise ValueError(f"{path}: not a classifier file")
    version, count, mlen = struct.unpack("<III", blob[8:20])
    if version != 1 or count != N_PARAMS:
        raise ValueError(f"{path}: version {version}, {count} parameters")
    info = json.loads(blob[20:20 + mlen].decode()) if mlen else {}
    params = np.frombuffer(blob[20 + mlen:20 + mlen + 4 * count], "<f4").astype(np.float32)
    if params.size != N_PARAMS:
        raise ValueError(f"{path}: truncated")
    return params, info
