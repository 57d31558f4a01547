"""The edge-proposal network (dr_mlp_*): matrix-core inference and GPU
training against a NumPy model of the same network."""
import ctypes

import numpy as np
import pytest

from delta_ray_amd import _abi

F, H = 24, 64
P_W1, P_B1 = 0, F * H
P_W2 = P_B1 + H
P_B2 = P_W2 + H * H
P_W3 = P_B2 + H
P_B3 = P_W3 + H


def bf16(a):
    """Round float32 to bfloat16 (nearest even), returned as float32."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u >> 16) & 1) + 0x7FFF
    return (((u + r) >> 16) << 16).astype(np.uint32).view(np.float32)


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[P_W1:P_B1].reshape(F, H), p[P_B1:P_W2], p[P_W2:P_B2].reshape(H, H), p[P_B2:P_W3], p[P_W3:P_B3],
            p[P_B3])


def forward_np(p, X, bf=False):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    if bf:
        W1, W2 = bf16(W1).astype(np.float64), bf16(W2).astype(np.float64)
        X = bf16(X).astype(np.float64)
    z1 = X @ W1 + b1
    h1 = np.maximum(z1, 0)
    if bf:
        h1 = bf16(h1).astype(np.float64)
    z2 = h1 @ W2 + b2
    h2 = np.maximum(z2, 0)
    return h2 @ w3 + b3, (X, z1, h1, z2, h2)


def grad_np(p, X, y, w, inv_total):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    o, (X, z1, h1, z2, h2) = forward_np(p, X)
    d = 2.0 * w * (o - y) * inv_total
    g = np.zeros(_abi.DR_MLP_PARAMS)
    g[P_W3:P_B3] = h2.T @ d
    g[P_B3] = d.sum()
    dz2 = np.outer(d, w3) * (z2 > 0)
    g[P_W2:P_B2] = (h1.T @ dz2).ravel()
    g[P_B2:P_W3] = dz2.sum(0)
    dz1 = (dz2 @ W2.T) * (z1 > 0)
    g[P_W1:P_B1] = (X.T @ dz1).ravel()
    g[P_B1:P_W2] = dz1.sum(0)
    return g, float(np.sum(w * (o - y) ** 2) * inv_total)


class Net:
    def __init__(self, seed):
        self.lib = _abi.load_library()
        h = ctypes.c_void_p()
        _abi.check(self.lib, self.lib.dr_mlp_create(seed, ctypes.byref(h)), "dr_mlp_create")
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.dr_mlp_destroy(self.h)

    def params(self, set_to=None):
        out = np.zeros(_abi.DR_MLP_PARAMS, np.float32)
        src = None if set_to is None else np.ascontiguousarray(set_to, np.float32)
        _abi.check(self.lib, self.lib.dr_mlp_params(self.h, out.ctypes.data, None if src is None else src.ctypes.data),
                   "dr_mlp_params")
        return out


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


@pytest.mark.gpu
def test_infer_matches_numpy(gpu):
    import torch
    net = Net(7)
    p = net.params()
    rng = np.random.default_rng(0)
    for n in (1, 17, 1000):
        X = rng.normal(size=(n, F)).astype(np.float32)
        xt = torch.from_numpy(X).to(gpu)
        out = torch.zeros(n, device=gpu)
        _abi.check(net.lib, net.lib.dr_mlp_infer(net.h, _ptr(xt), n, _ptr(out), None), "dr_mlp_infer")
        torch.cuda.synchronize()
        ref, _ = forward_np(p, X, bf=True)
        got = out.cpu().numpy()
        np.testing.assert_allclose(got, ref, rtol=2e-3, atol=2e-3 * np.abs(ref).max())


def _philox_indices(oracle, seed, step, nb, n):
    lib = oracle.load()
    out = np.zeros(nb, np.int64)
    key = (ctypes.c_uint32 * 2)(seed & 0xFFFFFFFF, seed >> 32)
    for i in range(nb):
        # counter of philox_init(seed, pass 3, sample = step, stream = i), block 0
        ctr = (ctypes.c_uint32 * 4)(step & 0xFFFFFFFF, ((step >> 32) ^ (3 << 24)) & 0xFFFFFFFF, i, 0)
        r = (ctypes.c_uint32 * 4)()
        lib.or_philox4x32(key, ctr, r)
        out[i] = ((int(r[0]) << 32) | int(r[1])) % n
    return out


@pytest.mark.gpu
def test_one_adam_step_matches_numpy(gpu, oracle):
    import torch
    net = Net(3)
    rng = np.random.default_rng(1)
    n, nb, lr, seed = 500, 128, 1e-3, 11
    X = rng.normal(size=(n, F)).astype(np.float32)
    y = rng.normal(size=n).astype(np.float32)
    w = rng.uniform(0.0, 2.0, size=n).astype(np.float32)
    p0 = net.params().astype(np.float64)
    idx = _philox_indices(oracle, seed, 0, nb, n)
    g, loss_ref = grad_np(p0, X[idx].astype(np.float64), y[idx], w[idx], 1.0 / nb)
    # Adam, first step: m = 0.1 g, v = 0.001 g^2, bias-corrected -> g / (|g| + eps)
    p1_ref = p0 - lr * g / (np.abs(g) + 1e-8)
    xt, yt, wt = (torch.from_numpy(a).to(gpu) for a in (X, y, w))
    loss = ctypes.c_float(0.0)
    _abi.check(net.lib, net.lib.dr_mlp_train(net.h, _ptr(xt), _ptr(yt), _ptr(wt), n, 1, nb, lr, seed,
                                             ctypes.byref(loss), None), "dr_mlp_train")
    p1 = net.params().astype(np.float64)
    assert abs(loss.value - loss_ref) <= 1e-4 * abs(loss_ref) + 1e-6
    big = np.abs(g) > 1e-4 * np.abs(g).max()   # tiny gradients: sign of g/|g| is rounding noise
    np.testing.assert_allclose(p1[big], p1_ref[big], rtol=1e-5, atol=1e-6)


@pytest.mark.gpu
def test_training_fits_a_smooth_target(gpu):
    import torch
    net = Net(5)
    rng = np.random.default_rng(2)
    n = 4096
    X = rng.normal(size=(n, F)).astype(np.float32)
    y = (0.8 * X[:, 0] - 0.5 * X[:, 3] + np.sin(X[:, 5]) + 0.3 * X[:, 7] * X[:, 8]).astype(np.float32)
    w = np.ones(n, np.float32)
    xt, yt, wt = (torch.from_numpy(a).to(gpu) for a in (X, y, w))
    loss0 = float(np.mean(y ** 2))
    loss = ctypes.c_float(0.0)
    _abi.check(net.lib, net.lib.dr_mlp_train(net.h, _ptr(xt), _ptr(yt), _ptr(wt), n, 600, 256, 3e-3, 9,
                                             ctypes.byref(loss), None), "dr_mlp_train")
    out = torch.zeros(n, device=gpu)
    _abi.check(net.lib, net.lib.dr_mlp_infer(net.h, _ptr(xt), n, _ptr(out), None), "dr_mlp_infer")
    torch.cuda.synchronize()
    mse = float(np.mean((out.cpu().numpy() - y) ** 2))
    assert mse < 0.1 * loss0, (mse, loss0)
