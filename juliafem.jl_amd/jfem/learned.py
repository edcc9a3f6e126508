"""Learned 3x3 block preconditioner: weights, training and evaluation.

The network (24 -> 64 -> 64 -> 6, GELU) runs on the GPU in the library's
MFMA kernels, forward and backward (libjfem_hip.so: jfm_learned_outputs,
jfm_learned_grad).  This module holds the FP32 master weights, packs them into
the BF16 blob the library consumes, and drives training:

* pairs: for seeded meshes of the cantilever family, the plain FP64
  block-Jacobi PCG is run to convergence (x*), then re-run for k iterations to
  give iterates x_k; each pair is (r_k = f - K x_k, e_k = x* - x_k);
* loss per pair: 1 - (z.r)^2 / ((z.Kz)(e.r)) with z = M r, the energy-norm
  error left after the best step along the preconditioned residual;
* optimiser: Adam on the FP32 master weights (a few thousand parameters, on
  the host); gradients come from the GPU kernels.

The last layer starts at zero, so the untrained network gives M = (1 + eps)
times block Jacobi.  There is no counterpart in the reference.
"""
from __future__ import annotations

import json
import os
import struct

import numpy as np

N_IN, N_H, N_OUT = 24, 64, 6
MAGIC = 0x4C4D464A
SHAPES = {"W1": (N_H, N_IN), "b1": (N_H,), "W2": (N_H, N_H), "b2": (N_H,), "W3": (N_OUT, N_H), "b3": (N_OUT,)}
ORDER = ("W1", "b1", "W2", "b2", "W3", "b3")     # gradient order of jfm_learned_grad
N_PARAMS = sum(int(np.prod(SHAPES[k])) for k in ORDER)
DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "learned_cantilever.bin")


def to_bf16(a: np.ndarray) -> np.ndarray:
    """FP32 -> BF16 bits, round to nearest even."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def from_bf16(h: np.ndarray) -> np.ndarray:
    return (np.asarray(h, dtype=np.uint32) << 16).view(np.float32)


def pack_blob(params: dict, shift, scale, eps: float) -> bytes:
    head = struct.pack("<6I", MAGIC, 1, N_IN, N_H, N_H, N_OUT)
    parts = [head]
    for k in ("W1", "W2", "W3"):
        parts.append(to_bf16(params[k]).tobytes())
    for k in ("b1", "b2", "b3"):
        parts.append(np.asarray(params[k], np.float32).tobytes())
    parts.append(np.asarray(shift, np.float32).tobytes())
    parts.append(np.asarray(scale, np.float32).tobytes())
    parts.append(struct.pack("<f", eps))
    return b"".join(parts)


def unpack_blob(blob: bytes) -> dict:
    magic, ver, ni, h1, h2, no = struct.unpack_from("<6I", blob, 0)
    if magic != MAGIC or (ni, h1, h2, no) != (N_IN, N_H, N_H, N_OUT):
        raise ValueError("not a 24-64-64-6 weight blob")
    off = 24
    out = {}
    for k in ("W1", "W2", "W3"):
        n = int(np.prod(SHAPES[k]))
        out[k] = np.frombuffer(blob, np.uint16, n, off).copy()
        off += 2 * n
    for k in ("b1", "b2", "b3"):
        n = SHAPES[k][0]
        out[k] = np.frombuffer(blob, np.float32, n, off).copy()
        off += 4 * n
    out["shift"] = np.frombuffer(blob, np.float32, N_IN, off).copy()
    off += 4 * N_IN
    out["scale"] = np.frombuffer(blob, np.float32, N_IN, off).copy()
    off += 4 * N_IN
    out["eps"] = struct.unpack_from("<f", blob, off)[0]
    return out


class LearnedPreconditioner:
    """FP32 master weights + feature normalisation + Adam state."""

    def __init__(self, seed: int = 0, eps: float = 1e-3):
        rng = np.random.default_rng(seed)
        self.params = {
            "W1": rng.standard_normal(SHAPES["W1"]).astype(np.float32) * np.float32(np.sqrt(2.0 / N_IN)),
            "b1": np.zeros(N_H, np.float32),
            "W2": rng.standard_normal(SHAPES["W2"]).astype(np.float32) * np.float32(np.sqrt(2.0 / N_H)),
            "b2": np.zeros(N_H, np.float32),
            "W3": np.zeros(SHAPES["W3"], np.float32),
            "b3": np.zeros(N_OUT, np.float32),
        }
        self.shift = np.zeros(N_IN, np.float32)
        self.scale = np.ones(N_IN, np.float32)
        self.eps = float(eps)
        self.m = {k: np.zeros_like(v) for k, v in self.params.items()}
        self.v = {k: np.zeros_like(v) for k, v in self.params.items()}
        self.t = 0
        self.history = []

    def blob(self) -> bytes:
        return pack_blob(self.params, self.shift, self.scale, self.eps)

    def fit_normalisation(self, feats):
        f = np.concatenate([np.asarray(x, np.float64) for x in feats])
        self.shift = f.mean(axis=0).astype(np.float32)
        sd = f.std(axis=0)
        self.scale = (1.0 / np.maximum(sd, 1e-6)).astype(np.float32)

    def flat_to_dict(self, g: np.ndarray) -> dict:
        out, off = {}, 0
        for k in ORDER:
            n = int(np.prod(SHAPES[k]))
            out[k] = np.asarray(g[off:off + n], np.float32).reshape(SHAPES[k])
            off += n
        return out

    def adam_step(self, grad: dict, lr: float = 1e-3, b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8):
        self.t += 1
        for k in ORDER:
            g = grad[k]
            self.m[k] = b1 * self.m[k] + (1 - b1) * g
            self.v[k] = b2 * self.v[k] + (1 - b2) * g * g
            mh = self.m[k] / (1 - b1 ** self.t)
            vh = self.v[k] / (1 - b2 ** self.t)
            self.params[k] = (self.params[k] - lr * mh / (np.sqrt(vh) + eps)).astype(np.float32)

    def save(self, path: str, meta: dict | None = None):
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        with open(path, "wb") as fh:
            fh.write(self.blob())
        with open(path + ".json", "w") as fh:
            json.dump({"format": "jfem learned preconditioner, 24-64-64-6 GELU, bf16 weights",
                       "bytes": len(self.blob()), "eps": self.eps, "adam_steps": self.t,
                       "history": self.history, **(meta or {})}, fh, indent=1)

    @classmethod
    def load(cls, path: str = DEFAULT_WEIGHTS) -> "LearnedPreconditioner":
        with open(path, "rb") as fh:
            blob = fh.read()
        w = unpack_blob(blob)
        lp = cls()
        for k in ("W1", "W2", "W3"):
            lp.params[k] = from_bf16(w[k]).reshape(SHAPES[k]).copy()
        for k in ("b1", "b2", "b3"):
            lp.params[k] = w[k].copy()
        lp.shift, lp.scale, lp.eps = w["shift"], w["scale"], float(w["eps"])
        return lp


def harvest_pairs(ctx, mesh, n_pairs: int = 8, rtol_star: float = 1e-10, seed: int = 0, coarse_nodes: int = 0):
    """(R, E): residual/error pairs from iterates of the FP64 block-Jacobi
    PCG (two-level with coarse_nodes > 0)."""
    xs, _, st = ctx.solve(method=0, op="matfree", rtol=rtol_star, coarse_nodes=coarse_nodes, raise_on_noconv=False)
    n_it = max(int(st["iterations"]), 4)
    rng = np.random.default_rng(seed)
    ks = np.unique(np.clip(np.geomspace(2, 0.98 * n_it, n_pairs).astype(int)
                           + rng.integers(0, 3, n_pairs), 1, n_it - 1))
    fixed = mesh.fixed_mask().astype(bool)
    R, E = [], []
    for k in ks:
        xk, _, _ = ctx.solve(method=0, op="matfree", rtol=1e-30, maxit=int(k), coarse_nodes=coarse_nodes,
                             raise_on_noconv=False)
        e = xs - xk
        e[fixed] = 0.0
        r = ctx.apply(e, op="matfree")     # = f - K x_k, consistent with e exactly
        R.append(r)
        E.append(e)
    return np.array(R), np.array(E)


def train(contexts, pair_sets, lp: LearnedPreconditioner, steps: int = 200, lr: float = 2e-3, log=None,
          coarse_nodes=0):
    """Adam on the mean loss over every (context, pairs) set; returns the loss
    history.  coarse_nodes > 0 (one value, or one per context) trains the
    blocks inside the two-level preconditioner."""
    for s in range(steps):
        blob = lp.blob()
        gsum = np.zeros(N_PARAMS)
        lsum = 0.0
        cns = coarse_nodes if isinstance(coarse_nodes, (list, tuple)) else [coarse_nodes] * len(contexts)
        for ctx, (R, E), cn in zip(contexts, pair_sets, cns):
            ctx.set_learned_weights(blob)
            loss, g = ctx.learned_grad(R, E, coarse_nodes=cn)
            gsum += g
            lsum += loss
        gsum /= len(contexts)
        lsum /= len(contexts)
        lp.adam_step(lp.flat_to_dict(gsum.astype(np.float32)), lr=lr)
        lp.history.append(lsum)
        if log and (s % 10 == 0 or s == steps - 1):
            log(f"step {s}: loss {lsum:.6f} |g| {np.linalg.norm(gsum):.3e}")
    return lp.history
