"""Numpy restatement of the orientation-proposal network -- test infrastructure only.

Forward, soft-target cross entropy and backward of the multilayer perceptron that
dragonfly_amd/proposal.py runs on the GPU (ReLU hidden layers, linear output, mean loss over
the batch), in float64, plus one Adam step. The GPU kernels are checked against it; it is
itself pinned by central finite differences in tests/test_oracle_mlp.py.
"""
import numpy as np


def forward(X, Ws, bs):
    """Activations [X, h1, ..., logits]; Ws[l] is (out, in)."""
    acts = [np.asarray(X, dtype='f8')]
    h = acts[0]
    for l, (W, b) in enumerate(zip(Ws, bs)):
        z = h @ np.asarray(W, 'f8').T + np.asarray(b, 'f8')
        if l < len(Ws) - 1:
            z = np.maximum(z, 0.)
        acts.append(z)
        h = z
    return acts


def softmax(Z):
    Z = Z - Z.max(1, keepdims=True)
    E = np.exp(Z)
    return E / E.sum(1, keepdims=True)


def loss_and_grads(X, Y, Ws, bs):
    """Mean over rows of -sum_j Y_j log softmax(Z)_j, and its gradients (gW, gb)."""
    acts = forward(X, Ws, bs)
    Z = acts[-1]
    B = Z.shape[0]
    P = softmax(Z)
    loss = float(-(Y * np.log(P)).sum() / B)
    g = (P - Y) / B
    gW, gb = [None] * len(Ws), [None] * len(Ws)
    for l in range(len(Ws) - 1, -1, -1):
        gW[l] = g.T @ acts[l]
        gb[l] = g.sum(0)
        if l > 0:
            g = (g @ np.asarray(Ws[l], 'f8')) * (acts[l] > 0)
    return loss, gW, gb


def adam(w, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8, wd=0.):
    """One Adam step (bias-corrected moments, decoupled from nothing: g + wd w)."""
    g = g + wd * w
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    w = w - lr * (m / (1 - b1 ** step)) / (np.sqrt(v / (1 - b2 ** step)) + eps)
    return w, m, v
