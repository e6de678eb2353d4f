"""CPU restatement of the learned warm start (features, network, loss and
gradients, node <-> layer resampling) in numpy.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the reference for the library's
clm_ws_* kernels; never by the library's product path.

The network mirrors the library's: BF16 inputs, weights and activations,
sums accumulated in (here) float64, GELU in its tanh form
z σ(2 sqrt(2/π) (z + 0.044715 z^3)), masked mean-square loss.  `bf16=False` drops every BF16 rounding, which gives the exact float64
function whose gradient finite differences pin.
"""
from __future__ import annotations

import math

import numpy as np

NODES = 32
SCALE_V = 1e-2
SCALE_T = 1.0


def dims(eh: bool):
    """(F, Fp, H, Nout)."""
    return (105, 128, 384, 64) if eh else (40, 64, 256, 32)


def to_bf16(x) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest, ties to even), as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u >> 16) & 1) + 0x7FFF
    return ((u + r) & 0xFFFF0000).astype(np.uint32).view(np.float32)


def layer_shapes(eh: bool):
    F, Fp, H, No = dims(eh)
    return [(H, Fp), (H, H), (H, H), (No, H)]


def unpack(params: np.ndarray, eh: bool):
    """Flat parameter vector -> [(Wt [N x K], b [N])] per layer."""
    out, o = [], 0
    for N, K in layer_shapes(eh):
        W = params[o:o + N * K].reshape(N, K).astype(np.float64)
        o += N * K
        b = params[o:o + N].astype(np.float64)
        o += N
        out.append((W, b))
    assert o == params.size
    return out


_C = 2.0 * math.sqrt(2.0 / math.pi)


def gelu(z):
    u = _C * (z + 0.044715 * z ** 3)
    return z / (1.0 + np.exp(-u))


def dgelu(z):
    u = _C * (z + 0.044715 * z ** 3)
    s = 1.0 / (1.0 + np.exp(-u))
    return s + z * s * (1.0 - s) * _C * (1.0 + 3 * 0.044715 * z * z)


def _r(x, bf16):
    return to_bf16(x).astype(np.float64) if bf16 else np.asarray(x, dtype=np.float64)


def forward(params, eh, X, bf16=True):
    """Returns Y and the per-layer (input, pre-activation) cache."""
    layers = unpack(params, eh)
    a = _r(X, bf16)
    cache = []
    for li, (W, b) in enumerate(layers):
        Wr = _r(W, bf16)
        z = a @ Wr.T + b
        if bf16:
            z = z.astype(np.float32).astype(np.float64)
        cache.append((a, z))
        if li < 3:
            a = _r(gelu(z), bf16)
        else:
            a = z
    return a, cache


def loss_and_grad(params, eh, X, T, mask, bf16=True):
    """Masked MSE sum(mask (Y - T)^2) / (M Nout) and its flat gradient."""
    layers = unpack(params, eh)
    Y, cache = forward(params, eh, X, bf16)
    M, No = Y.shape
    d = Y - T
    loss = float((mask[:, None] * d * d).sum() / (M * No))
    g = _r(2.0 * mask[:, None] * d / (M * No), bf16)
    grads = [None] * 4
    for li in range(3, -1, -1):
        a, z = cache[li]
        if li < 3:
            g = _r(g * dgelu(z), bf16)
        grads[li] = (g.T @ a, g.sum(axis=0))
        W = _r(layers[li][0], bf16)
        g = g @ W
    flat = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return loss, flat


# ---- resampling and features (restating the library's feature kernels) ----

def at_nodes(v: np.ndarray) -> np.ndarray:
    """Cell values (level 0 = bottom) -> 32 node values, linear in layer index."""
    N = v.size
    q = np.clip((np.arange(NODES) + 0.5) * N / NODES - 0.5, 0.0, N - 1)
    i0 = q.astype(int)
    i1 = np.minimum(i0 + 1, N - 1)
    t = q - i0
    return (1 - t) * v[i0] + t * v[i1]


def at_cells(o: np.ndarray, N: int) -> np.ndarray:
    """32 node values -> N cell values."""
    q = np.clip((np.arange(N) + 0.5) * NODES / N - 0.5, 0.0, NODES - 1)
    j0 = q.astype(int)
    j1 = np.minimum(j0 + 1, NODES - 1)
    t = q - j0
    return (1 - t) * o[j0] + t * o[j1]


def features(batch, state, top_water, top_heat, dt, consts):
    """Feature rows [ncol x Fp] of a CSR state (float64, before BF16 rounding)."""
    eh = batch.model == 1
    F, Fp, H, No = dims(eh)
    P = batch.params
    X = np.zeros((batch.ncol, Fp))
    for c in range(batch.ncol):
        a, b = batch.col_ptr[c], batch.col_ptr[c + 1]
        N = b - a
        nu, thr, al, n, ks = P[0, c], P[1, c], P[2, c], P[3, c], P[4, c]
        v = state["vartheta_l"][a:b]
        row = [(at_nodes(v) - thr) / (nu - thr)]
        if eh:
            ti, re = state["theta_i"][a:b], state["rho_e_int"][a:b]
            thl = np.minimum(nu - ti, v)
            rcs = P[12, c] + thl * consts["RHO_L"] * consts["CP_L"] + \
                ti * consts["RHO_I"] * consts["CP_I"]
            T = consts["T_0"] + (re + ti * consts["RHO_I"] * consts["LH_F0"]) / rcs
            row += [(at_nodes(T) - 273.15) / 10.0, at_nodes(ti) / nu]
        tail = [math.log10(al), n - 2.0, (math.log10(ks) + 5.5) / 1.5, (nu - 0.45) / 0.1,
                (thr - 0.06) / 0.04, top_water[c] / 1e-5]
        if eh:
            tail.append(top_heat[c] / 100.0)
        tail += [dt / 3600.0, N / 40.0]
        r = np.concatenate(row + [np.array(tail)])
        assert r.size == F
        # non-finite inputs enter as 0, finite ones are clamped to [-20, 20]
        X[c, :F] = np.clip(np.where(np.isfinite(r), r, 0.0), -20.0, 20.0)
    return X
