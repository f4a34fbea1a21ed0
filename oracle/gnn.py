"""CPU oracle of the learned vertex-block preconditioner (numpy, FP64).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the HIP
training and inference kernels.  The product path never imports it.

The reference has no learned preconditioner (it defers one:
REVIEW_AND_IMPROVEMENT_PLAN.md:377-379), so this module restates the
project's own definition.  Its backward pass is pinned by central finite
differences in tests/test_gnn_oracle.py; its preconditioner apply is pinned
by the closed form (all coefficients one and no network = truncated Neumann
series of the Jacobi splitting, which tends to A^-1).

Definitions (b components, K = degree, width 32, 3 message-passing layers):

  features   rho_ic = sum_e |a_ec| / d_ic,
             x_i    = [rho_ic, log(max(1 - rho_ic, 1e-6)) / 8,
                       log(deg_i max_e |a_ec| / sum_e |a_ec|) / 2   (c = 0..b-1),
                       deg_i / 8]
             (the third group is the row's edge anisotropy: 0 when every edge
             of the row carries the same coefficient)
  edge weight w_e   = (1/b) sum_c |a_ec| / d_ic            (e in row i)
  encoder    h0 = tanh(W_in x + b_in)
  layer l    agg_i = sum_e w_e h_{col e},  h = tanh(W_s h + W_n agg + b)
  decoder    o = W_out h3 + b_out,  index k*b + c
  coefficients gamma_kic = g_kc + o_kci   (additive, unbounded: W_out = b_out = 0
             gives the global polynomial g alone)
  apply      u_0 = D^-1 r,  u_(k+1) = -D^-1 A_off u_k,  z = sum_k gamma_k * u_k
  loss       L = (1/P) sum_p |D^-1/2 A (z_p - e_p)|^2 / |D^-1/2 A e_p|^2
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

WIDTH = 32
LAYERS = 3


def feature_count(b: int) -> int:
    return 3 * b + 1


@dataclass
class Layout:
    """Offsets of each parameter block inside the flat parameter vector
    (the HIP model uses the same order)."""
    b: int
    degree: int

    def blocks(self):
        f0, w, o = feature_count(self.b), WIDTH, (self.degree + 1) * self.b
        out = [("w_in", (w, f0)), ("b_in", (w,))]
        for layer in range(LAYERS):
            out += [(f"ws{layer}", (w, w)), (f"wn{layer}", (w, w)), (f"b{layer}", (w,))]
        out += [("w_out", (o, w)), ("b_out", (o,)), ("g", (o,))]
        return out

    def count(self) -> int:
        return sum(int(np.prod(s)) for _, s in self.blocks())

    def unpack(self, flat):
        out, pos = {}, 0
        for name, shape in self.blocks():
            size = int(np.prod(shape))
            out[name] = np.asarray(flat[pos:pos + size], np.float64).reshape(shape)
            pos += size
        return out

    def pack(self, params) -> np.ndarray:
        return np.concatenate([np.asarray(params[name], np.float64).ravel() for name, _ in self.blocks()])


def bf16_round(a) -> np.ndarray:
    """Round FP32 values to bfloat16 (nearest even), returned as float64."""
    f = np.ascontiguousarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


class Graph:
    """Split-view system (rows = vertices) for the oracle."""

    def __init__(self, nodes, b, offsets, columns, diagonal, off_diagonal):
        self.n, self.b = int(nodes), int(b)
        self.offsets = np.asarray(offsets, np.int64)
        self.columns = np.asarray(columns, np.int64)
        self.diag = np.asarray(diagonal, np.float64).reshape(self.n, self.b)
        self.off = np.asarray(off_diagonal, np.float64).reshape(-1, self.b)
        self.rows = np.repeat(np.arange(self.n), np.diff(self.offsets))
        self.deg = np.diff(self.offsets).astype(np.float64)
        # mirror[e]: entry of (col e, row e)
        key = self.rows * self.n + self.columns
        rkey = self.columns * self.n + self.rows
        order = np.argsort(key)
        pos = np.searchsorted(key[order], rkey)
        if np.any(pos >= key.size) or np.any(key[order][np.minimum(pos, key.size - 1)] != rkey):
            raise ValueError("pattern is not structurally symmetric")
        self.mirror = order[pos]

    # A_off x per component (x: n x b)
    def off_apply(self, x):
        y = np.zeros((self.n, self.b))
        np.add.at(y, self.rows, self.off * x[self.columns])
        return y

    def off_apply_t(self, x):
        y = np.zeros((self.n, self.b))
        np.add.at(y, self.rows, self.off[self.mirror] * x[self.columns])
        return y

    def apply(self, x):
        return self.diag * x + self.off_apply(x)

    def apply_t(self, x):
        return self.diag * x + self.off_apply_t(x)

    def aggregate(self, w, h):
        y = np.zeros((self.n, h.shape[1]))
        np.add.at(y, self.rows, w[:, None] * h[self.columns])
        return y

    def aggregate_t(self, w, h):
        y = np.zeros((self.n, h.shape[1]))
        np.add.at(y, self.rows, w[self.mirror][:, None] * h[self.columns])
        return y


def features(g: Graph):
    absoff = np.abs(g.off)
    s = np.zeros((g.n, g.b))
    np.add.at(s, g.rows, absoff)
    mx = np.zeros((g.n, g.b))
    np.maximum.at(mx, g.rows, absoff)
    rho = s / g.diag
    lg = np.log(np.maximum(1.0 - rho, 1e-6)) / 8.0
    deg = np.maximum(g.deg, 1.0)[:, None]
    aniso = np.log(np.maximum(deg * mx / np.maximum(s, 1e-300), 1.0)) / 2.0
    x = np.concatenate([rho, lg, aniso, (g.deg / 8.0)[:, None]], axis=1)
    w = (absoff / g.diag[g.rows]).mean(axis=1)
    return x, w


def forward(g: Graph, p: dict, degree: int):
    """Returns gamma (degree+1, n, b) and the cache for backward."""
    x, w = features(g)
    hs = [np.tanh(x @ p["w_in"].T + p["b_in"])]
    aggs = []
    for layer in range(LAYERS):
        agg = g.aggregate(w, hs[-1])
        aggs.append(agg)
        hs.append(np.tanh(hs[-1] @ p[f"ws{layer}"].T + agg @ p[f"wn{layer}"].T + p[f"b{layer}"]))
    o = hs[-1] @ p["w_out"].T + p["b_out"]
    k1 = degree + 1
    gam = (p["g"][None, :] + o).reshape(g.n, k1, g.b).transpose(1, 0, 2)
    return gam, dict(x=x, w=w, hs=hs, aggs=aggs)


def backward(g: Graph, p: dict, degree: int, cache: dict, dgamma):
    """Gradient of a scalar with respect to the parameters, given d/dgamma."""
    k1 = degree + 1
    dg_flat = np.asarray(dgamma).transpose(1, 0, 2).reshape(g.n, k1 * g.b)
    grads = {"g": dg_flat.sum(0)}
    do = dg_flat
    hs, aggs, w = cache["hs"], cache["aggs"], cache["w"]
    grads["w_out"] = do.T @ hs[-1]
    grads["b_out"] = do.sum(0)
    dh = do @ p["w_out"]
    for layer in reversed(range(LAYERS)):
        h = hs[layer + 1]
        du = dh * (1.0 - h * h)
        grads[f"ws{layer}"] = du.T @ hs[layer]
        grads[f"wn{layer}"] = du.T @ aggs[layer]
        grads[f"b{layer}"] = du.sum(0)
        dh = du @ p[f"ws{layer}"] + g.aggregate_t(w, du @ p[f"wn{layer}"])
    du = dh * (1.0 - hs[0] * hs[0])
    grads["w_in"] = du.T @ cache["x"]
    grads["b_in"] = du.sum(0)
    return grads


def basis(g: Graph, r, degree: int):
    """u_0 = D^-1 r, u_(k+1) = -D^-1 A_off u_k; r is (n, b)."""
    u = [r / g.diag]
    for _ in range(degree):
        u.append(-g.off_apply(u[-1]) / g.diag)
    return np.array(u)


def apply(g: Graph, gamma, r):
    u = basis(g, r, gamma.shape[0] - 1)
    return (gamma * u).sum(0)


def loss_and_dgamma(g: Graph, gamma, residuals, errors):
    """Loss over the pairs and d loss / d gamma."""
    degree = gamma.shape[0] - 1
    sq = 1.0 / np.sqrt(g.diag)
    P = len(residuals)
    loss = 0.0
    dgam = np.zeros_like(gamma)
    for r, e in zip(residuals, errors):
        u = basis(g, r, degree)
        z = (gamma * u).sum(0)
        target = sq * g.apply(e)
        norm = float((target * target).sum())
        y = sq * g.apply(z) - target
        loss += float((y * y).sum()) / norm / P
        gz = (2.0 / (P * norm)) * g.apply_t(sq * y)
        dgam += gz[None] * u
    return loss, dgam


def loss_and_grad(g: Graph, flat, degree: int, residuals, errors):
    lay = Layout(g.b, degree)
    p = lay.unpack(flat)
    gam, cache = forward(g, p, degree)
    loss, dgam = loss_and_dgamma(g, gam, residuals, errors)
    return loss, lay.pack(backward(g, p, degree, cache, dgam))


def init_parameters(b: int, degree: int, seed: int = 0) -> np.ndarray:
    """Glorot-uniform weights, zero biases, zero decoder (so gamma = g at
    start) and g = 1 (truncated Neumann series)."""
    lay = Layout(b, degree)
    rng = np.random.default_rng(seed)
    p = {}
    for name, shape in lay.blocks():
        if len(shape) == 2 and name != "w_out":
            lim = np.sqrt(6.0 / (shape[0] + shape[1]))
            p[name] = rng.uniform(-lim, lim, shape)
        elif name == "g":
            p[name] = np.ones(shape)
        else:
            p[name] = np.zeros(shape)
    return lay.pack(p)
