"""ORACLE — test infrastructure only (see oracle/__init__.py).

CPU restatement of the kernel surrogate's dense network, in float64, with the
GPU's one deliberate rounding made explicit: every matrix-product operand is
rounded to bfloat16 (round to nearest even) before the product, as the GPU's
bf16 matrix-core GEMMs do; all sums are exact float64 here (FP32 on the GPU).

Network: hidden layers z = bf(a) @ bf(W) + b, a = tanh(z); last layer linear;
loss = mean((y - t)^2).  Gradients by the chain rule with the same roundings
(dW = bf(a_prev)^T @ bf(dz), da_prev = bf(dz) @ bf(W)^T, db = sum of dz).
The surrogate has no counterpart in the reference; this module checks the
GPU forward/backward kernels (tests/test_gpu_surrogate.py) and is itself
pinned by central finite differences (tests/test_oracle_surrogate.py).
"""
import numpy as np


def bf16(x):
    """Round to bfloat16 (nearest even), returned as float64."""
    f = np.ascontiguousarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def forward(params, x, round_bf16=True):
    """params: list of (W [in, out], b [out]).  Returns the activations
    [x, a_1, ..., a_L] (a_L linear)."""
    rd = bf16 if round_bf16 else (lambda v: np.asarray(v, dtype=np.float64))
    acts = [np.asarray(x, dtype=np.float64)]
    for i, (W, b) in enumerate(params):
        z = rd(acts[-1]) @ rd(W) + np.asarray(b, dtype=np.float64)
        acts.append(np.tanh(z) if i < len(params) - 1 else z)
    return acts


def loss_and_grads(params, x, t, round_bf16=True):
    """Mean squared error and its gradients [(dW, db), ...]."""
    rd = bf16 if round_bf16 else (lambda v: np.asarray(v, dtype=np.float64))
    acts = forward(params, x, round_bf16)
    y = acts[-1]
    t = np.asarray(t, dtype=np.float64)
    loss = float(np.mean((y - t) ** 2))
    dz = 2.0 * (y - t) / y.size
    grads = [None] * len(params)
    for i in range(len(params) - 1, -1, -1):
        W, _ = params[i]
        a_prev = acts[i]
        grads[i] = (rd(a_prev).T @ rd(dz), dz.sum(axis=0))
        if i > 0:
            da = rd(dz) @ rd(W).T
            dz = da * (1.0 - acts[i] ** 2)
    return loss, grads
