"""The numpy restatement of the warm-start network (tests/mlp_ref.py) against
central finite differences, so that the GPU gradient check compares with a
gradient that is itself pinned."""
import numpy as np

import mlp_ref


def test_parameter_count():
    n = sum(a * b + b for a, b in zip(mlp_ref.DIMS[:-1], mlp_ref.DIMS[1:]))
    assert n == mlp_ref.NPARAM == 37007


def test_reference_gradient_finite_differences():
    rng = np.random.default_rng(0)
    p = rng.normal(0.0, 0.2, mlp_ref.NPARAM)
    x = rng.normal(size=(7, 15))
    t = rng.normal(size=(7, 15))
    loss, g = mlp_ref.loss_and_grad(p, x, t)
    idx = rng.choice(mlp_ref.NPARAM, 40, replace=False)
    idx = np.concatenate([idx, [0, 15 * 128, 15 * 128 + 128, mlp_ref.NPARAM - 1]])   # W1, b1, W2, b4 corners
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (mlp_ref.loss_and_grad(pp, x, t)[0] - mlp_ref.loss_and_grad(pm, x, t)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-6 * abs(fd), (i, fd, g[i])


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -3.14159265, 1e-30], dtype=np.float32)
    r = mlp_ref.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0            # half-way, even
    assert r[2] == 1.0 + 2 ** -7                   # above half-way
    assert abs(r[3] + 3.140625) < 1e-12
    assert abs(r[4] - 1e-30) / 1e-30 < 2 ** -8
