"""The network oracle (oracle/mlp.py) pinned by central finite differences of its own loss
on a small float64 network, and the Adam step against its closed form on one coordinate."""
import numpy as np

from oracle import mlp


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    dims = [7, 6, 5, 4]
    Ws = [rng.normal(size=(dims[i + 1], dims[i])) for i in range(3)]
    bs = [rng.normal(size=dims[i + 1]) * 0.1 for i in range(3)]
    X = rng.normal(size=(5, dims[0]))
    Y = rng.random((5, dims[-1]))
    Y /= Y.sum(1, keepdims=True)
    loss, gW, gb = mlp.loss_and_grads(X, Y, Ws, bs)
    h = 1e-6
    for l in range(3):
        for (i, j) in ((0, 0), (dims[l + 1] - 1, dims[l] - 1), (1, 2)):
            Wp = [w.copy() for w in Ws]
            Wm = [w.copy() for w in Ws]
            Wp[l][i, j] += h
            Wm[l][i, j] -= h
            fd = (mlp.loss_and_grads(X, Y, Wp, bs)[0] - mlp.loss_and_grads(X, Y, Wm, bs)[0]) / (2 * h)
            assert abs(fd - gW[l][i, j]) < 1e-6 * max(1., abs(fd)), (l, i, j)
        bp = [b.copy() for b in bs]
        bm = [b.copy() for b in bs]
        bp[l][1] += h
        bm[l][1] -= h
        fd = (mlp.loss_and_grads(X, Y, Ws, bp)[0] - mlp.loss_and_grads(X, Y, Ws, bm)[0]) / (2 * h)
        assert abs(fd - gb[l][1]) < 1e-6 * max(1., abs(fd))
    assert loss > 0


def test_adam_first_step_is_sign_step():
    # with zero moments, the first bias-corrected step is lr * g / (|g| + eps)
    w, m, v = mlp.adam(np.array([1., 2.]), np.array([0.5, -3.]), np.zeros(2), np.zeros(2), 1, 0.1)
    assert np.allclose(w, [1. - 0.1, 2. + 0.1], atol=1e-6)
