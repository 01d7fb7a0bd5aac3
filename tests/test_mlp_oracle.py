"""CPU pins of the MLP oracle: analytic backward against central finite differences,
and the closed-form softmax/cross-entropy gradient on a one-layer special case."""
import numpy as np

import mlp_ref


def _params(rng, width, hidden, classes):
    return [rng.standard_normal((hidden, width)) * 0.3, rng.standard_normal(hidden) * 0.1,
            rng.standard_normal((hidden, hidden)) * 0.3, rng.standard_normal(hidden) * 0.1,
            rng.standard_normal((classes + 1, hidden)) * 0.3, rng.standard_normal(classes + 1) * 0.1]


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    width, hidden, classes, B = 6, 5, 7, 4
    params = _params(rng, width, hidden, classes)
    X = rng.standard_normal((B, width))
    labels = rng.integers(0, classes, B)
    stop = rng.standard_normal(B)
    w = 0.3
    _, _, grads = mlp_ref.loss_grad(params, X, labels, stop, w, emulate_bf16=False)
    eps = 1e-6
    for t, (P, G) in enumerate(zip(params, grads)):
        flat = P.reshape(-1)
        for i in rng.choice(flat.size, size=min(6, flat.size), replace=False):
            old = flat[i]
            flat[i] = old + eps
            lp = mlp_ref.total_loss(params, X, labels, stop, w)
            flat[i] = old - eps
            lm = mlp_ref.total_loss(params, X, labels, stop, w)
            flat[i] = old
            num = (lp - lm) / (2 * eps)
            assert abs(num - G.reshape(-1)[i]) <= 1e-6 * max(1.0, abs(num)), (t, i, num, G.reshape(-1)[i])


def test_softmax_gradient_closed_form():
    # all-zero last layer: uniform softmax, so dZ = (1/C - onehot) / B exactly
    rng = np.random.default_rng(1)
    width, hidden, classes, B = 4, 3, 5, 2
    params = _params(rng, width, hidden, classes)
    params[4][:] = 0.0
    params[5][:] = 0.0
    X = rng.standard_normal((B, width))
    labels = np.array([1, 3])
    xent, mse, grads = mlp_ref.loss_grad(params, X, labels, np.zeros(B), 0.0, emulate_bf16=False)
    assert abs(xent - np.log(classes)) < 1e-12
    expect_db3 = np.full(classes + 1, B / classes / B)
    expect_db3[-1] = 0.0
    for lab in labels:
        expect_db3[lab] -= 1.0 / B
    assert np.allclose(grads[5], expect_db3, atol=1e-15)


def test_top2_conf():
    Z = np.array([[1.0, 3.0, 3.0, 0.0, 9.0]])
    order, conf, stop = mlp_ref.top2_conf(Z)
    assert order.tolist() == [[1, 2]]
    p = np.exp(Z[0, :4] - 3) / np.exp(Z[0, :4] - 3).sum()
    assert abs(conf[0] - (p[1] + p[2])) < 1e-15 and stop[0] == 9.0
