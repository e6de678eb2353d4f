"""CPU checks of the warm-start network restatement (oracle/mlp_oracle.py):
BF16 rounding and GELU against torch's own implementations, the float64
gradient against torch autograd and finite differences, and the node/layer
resampling against exact cases."""
import numpy as np
import pytest
import torch

from oracle import mlp_oracle as mo


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(4000).astype(np.float32) * 10.0 ** rng.integers(-8, 8, 4000),
                        np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, -1 - 2 ** -8, 0.0],
                                 dtype=np.float32)]).astype(np.float32)
    ours = mo.to_bf16(x)
    ref = torch.from_numpy(x).to(torch.bfloat16).float().numpy()
    assert np.array_equal(ours, ref)
    assert ours[-4] == 1.0 and ours[-3] == 1 + 2 ** -6  # ties to even


def test_gelu_matches_torch_and_derivative():
    z = np.linspace(-8, 8, 2001)
    ref = torch.nn.functional.gelu(torch.from_numpy(z), approximate="tanh").numpy()
    assert np.abs(mo.gelu(z) - ref).max() <= 1e-14
    h = 1e-6
    fd = (mo.gelu(z + h) - mo.gelu(z - h)) / (2 * h)
    assert np.abs(mo.dgelu(z) - fd).max() <= 1e-8


@pytest.mark.parametrize("eh", [False, True])
def test_exact_gradient_matches_torch_autograd(eh):
    rng = np.random.default_rng(1 + eh)
    F, Fp, H, No = mo.dims(eh)
    P = sum(n * k + n for n, k in mo.layer_shapes(eh))
    params = (rng.standard_normal(P) * 0.05).astype(np.float32)
    M = 16
    X = np.zeros((M, Fp))
    X[:, :F] = rng.standard_normal((M, F))
    T = rng.standard_normal((M, No))
    mask = (rng.random(M) > 0.3).astype(np.float64)
    loss, g = mo.loss_and_grad(params, eh, X, T, mask, bf16=False)
    # the same network in torch float64 with autograd
    tp = torch.from_numpy(params.astype(np.float64)).requires_grad_(True)
    o, a = 0, torch.from_numpy(X)
    for li, (n, k) in enumerate(mo.layer_shapes(eh)):
        W = tp[o:o + n * k].reshape(n, k)
        o += n * k
        b = tp[o:o + n]
        o += n
        a = a @ W.T + b
        if li < 3:
            a = torch.nn.functional.gelu(a, approximate="tanh")
    tl = ((torch.from_numpy(mask)[:, None] * (a - torch.from_numpy(T)) ** 2).sum()) / (M * No)
    tl.backward()
    assert abs(loss - tl.item()) <= 1e-12 * abs(loss)
    assert np.abs(g - tp.grad.numpy()).max() <= 1e-10 * np.abs(g).max()
    # and a few finite differences
    for i in rng.choice(P, 8, replace=False):
        h = 1e-3
        p1, p2 = params.astype(np.float64), params.astype(np.float64)
        p1[i] += h
        p2[i] -= h
        fd = (mo.loss_and_grad(p1, eh, X, T, mask, bf16=False)[0] -
              mo.loss_and_grad(p2, eh, X, T, mask, bf16=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(abs(g).max(), 1e-12), i


def test_resampling_exact_cases():
    v = np.random.default_rng(3).standard_normal(32)
    assert np.allclose(mo.at_cells(mo.at_nodes(v), 32), v, rtol=0, atol=1e-15)
    for N in (32, 40, 77, 120):  # linear in layer index: nodes are exact
        lin = 0.3 + 0.01 * np.arange(N)
        q = (np.arange(32) + 0.5) * N / 32 - 0.5
        assert np.allclose(mo.at_nodes(lin), 0.3 + 0.01 * q, atol=1e-14)
    for N in (10, 15, 25):  # constant profiles survive any resampling
        assert np.allclose(mo.at_cells(mo.at_nodes(np.full(N, 0.2)), N), 0.2, atol=1e-15)
