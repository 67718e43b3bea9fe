"""Pins the MLP oracle: its FP64 gradients equal central finite differences,
and BF16 rounding matches the IEEE round-to-nearest-even definition."""
import numpy as np

from oracle.mlp_ref import bf16_round, loss_and_grad


def test_bf16_rounding_known_values():
    # 1 + 2^-8 lies between bf16 neighbours 1 and 1 + 2^-7: ties go to even (1.0)
    assert bf16_round(1 + 2**-8) == 1.0
    assert bf16_round(1 + 3 * 2**-8) == 1 + 2**-6  # rounds up to the even neighbour
    assert bf16_round(-2.5) == -2.5
    assert bf16_round(1 / 3) == 0.333984375


def test_fp64_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    dims = [7, 5, 5, 1]
    n = sum(a * b + b for a, b in zip(dims[:-1], dims[1:]))
    params = rng.normal(0, 0.5, n)
    x = rng.normal(0, 1, (9, 7))
    t = rng.normal(0, 1, 9)
    loss, g = loss_and_grad(params, dims, x, t, bf16=False)
    for k in rng.choice(n, 25, replace=False):
        e = np.zeros(n)
        e[k] = 1e-6
        lp, _ = loss_and_grad(params + e, dims, x, t, bf16=False)
        lm, _ = loss_and_grad(params - e, dims, x, t, bf16=False)
        fd = (lp - lm) / 2e-6
        assert abs(fd - g[k]) < 1e-6 * (1 + abs(fd)), (k, fd, g[k])
