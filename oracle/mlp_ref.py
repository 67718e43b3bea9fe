"""CPU oracle of the MLP surrogate (test infrastructure only).

numpy FP64 restatement of qcsim_amd/csrc/mlp.hip: in -> hidden (ReLU) x L ->
1, MSE loss. With `bf16=True` it rounds exactly where the HIP code rounds
(inputs, weights and hidden activations to BF16, round-to-nearest-even on
the top 16 bits of the FP32 value; biases stay FP32), so its gradients are
the GPU's up to FP32-vs-FP64 accumulation order. With `bf16=False` it is the
plain FP64 network, whose gradients the tests check against central finite
differences.
"""
import numpy as np


def bf16_round(x):
    """float -> bf16 -> float, round to nearest even on the FP32 bit pattern."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(params, dims):
    """Flat FP32 vector -> [(W (in x out), b (out)), ...]."""
    out, off = [], 0
    for a, b in zip(dims[:-1], dims[1:]):
        W = params[off: off + a * b].reshape(a, b)
        off += a * b
        bias = params[off: off + b]
        off += b
        out.append((np.asarray(W, dtype=np.float64), np.asarray(bias, dtype=np.float64)))
    return out


def forward(params, dims, x, bf16=True):
    r = bf16_round if bf16 else (lambda v: np.asarray(v, dtype=np.float64))
    layers = unpack(params, dims)
    a = r(x)
    acts = [a]
    for W, b in layers[:-1]:
        z = a @ r(W) + b
        a = r(np.maximum(z, 0.0))
        acts.append(a)
    W, b = layers[-1]
    y = (a @ r(W) + b)[:, 0]
    return y, acts


def loss_and_grad(params, dims, x, t, bf16=True):
    r = bf16_round if bf16 else (lambda v: np.asarray(v, dtype=np.float64))
    layers = unpack(params, dims)
    y, acts = forward(params, dims, x, bf16)
    B = x.shape[0]
    d = y - t
    loss = float(np.mean(d * d))
    dz = (2.0 * d / B)[:, None]
    grads = []
    for l in range(len(layers) - 1, -1, -1):
        W, b = layers[l]
        a_in = acts[l]
        gW = a_in.T @ dz
        gb = dz.sum(axis=0)
        grads.append((gW, gb))
        if l == 0:
            break
        dz = (dz @ r(W).T) * (acts[l] > 0)
    flat = []
    for gW, gb in reversed(grads):
        flat.append(gW.reshape(-1))
        flat.append(gb)
    return loss, np.concatenate(flat)
