"""MLP HIP kernels against the CPU oracle: forward outputs and flat
gradients on the same BF16-rounded weights and inputs; training lowers the
loss."""
import ctypes as C

import numpy as np
import pytest

from oracle.mlp_ref import forward, loss_and_grad
from qcsim_amd._lib import check, lib

pytestmark = pytest.mark.gpu
_PF = C.POINTER(C.c_float)


def make(in_dim, hidden, layers, seed=1):
    h = C.c_void_p()
    check(lib().qcs_mlp_create(in_dim, hidden, layers, seed, C.byref(h)))
    n = lib().qcs_mlp_num_params(h)
    p = np.zeros(n, dtype=np.float32)
    check(lib().qcs_mlp_get_params(h, p.ctypes.data_as(_PF)))
    return h, p, [in_dim] + [hidden] * layers + [1]


@pytest.mark.parametrize("in_dim,hidden,layers,batch", [(32, 256, 3, 256), (17, 64, 2, 100), (8, 16, 1, 7)])
def test_forward_and_gradients_match_oracle(in_dim, hidden, layers, batch):
    h, p, dims = make(in_dim, hidden, layers)
    rng = np.random.default_rng(in_dim)
    # non-zero biases exercise the bias paths
    p = p + rng.normal(0, 0.05, p.shape).astype(np.float32)
    check(lib().qcs_mlp_set_params(h, p.ctypes.data_as(_PF)))
    x = rng.normal(0, 1, (batch, in_dim)).astype(np.float32)
    t = rng.normal(0, 1, batch).astype(np.float32)
    y = np.zeros(batch, dtype=np.float32)
    check(lib().qcs_mlp_forward(h, x.ctypes.data_as(_PF), batch, y.ctypes.data_as(_PF)))
    y_ref, _ = forward(p.astype(np.float64), dims, x)
    assert np.max(np.abs(y - y_ref)) < 2e-3 * (1 + np.max(np.abs(y_ref)))
    g = np.zeros(p.shape[0], dtype=np.float32)
    loss = C.c_float()
    check(lib().qcs_mlp_gradients(h, x.ctypes.data_as(_PF), t.ctypes.data_as(_PF), batch, g.ctypes.data_as(_PF),
                                  C.byref(loss)))
    l_ref, g_ref = loss_and_grad(p.astype(np.float64), dims, x, t)
    assert abs(loss.value - l_ref) < 2e-3 * (1 + l_ref)
    rel = np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref)
    assert rel < 1e-2, rel
    lib().qcs_mlp_destroy(h)


def test_training_reduces_loss():
    h, p, dims = make(16, 128, 3)
    rng = np.random.default_rng(3)
    X = rng.normal(0, 1, (8192, 16)).astype(np.float32)
    Y = (np.sin(X[:, 0]) + 0.5 * X[:, 1] * X[:, 2]).astype(np.float32)
    losses = np.zeros(8, dtype=np.float32)
    check(lib().qcs_mlp_train(h, X.ctypes.data_as(_PF), Y.ctypes.data_as(_PF), X.shape[0], 256, 8, 2e-3, 7,
                              losses.ctypes.data_as(_PF)))
    assert losses[-1] < 0.25 * losses[0], losses
    lib().qcs_mlp_destroy(h)
