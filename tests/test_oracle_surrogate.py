"""Pins of the surrogate network's oracle (CPU): analytic gradients against
central finite differences, and the bf16 rounding against hand cases."""
import numpy as np

from oracle import mlp


def test_bf16_rounding_hand_cases():
    # 1 + 2^-8 is halfway between bf16 neighbours 1 and 1 + 2^-7: ties to even (1)
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.0e38, 1e-40])
    r = mlp.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0
    assert r[2] == 1.0 + 2 ** -6          # 1 + 3*2^-8 rounds up to 1 + 4*2^-8
    assert r[3] == -2.5
    assert np.isfinite(r[4])
    assert abs(r[5] - 1e-40) < 1e-41


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    dims = [5, 7, 6, 3]
    params = [(rng.normal(0, 0.6, (dims[i], dims[i + 1])), rng.normal(0, 0.2, dims[i + 1]))
              for i in range(3)]
    x = rng.normal(size=(11, 5))
    t = rng.normal(size=(11, 3))
    loss, grads = mlp.loss_and_grads(params, x, t, round_bf16=False)
    h = 1e-6
    for li, (W, b) in enumerate(params):
        for arr, g in ((W, grads[li][0]), (b, grads[li][1])):
            flat = arr.reshape(-1)
            gf = np.asarray(g).reshape(-1)
            for k in range(0, flat.size, max(1, flat.size // 7)):
                old = flat[k]
                flat[k] = old + h
                lp, _ = mlp.loss_and_grads(params, x, t, round_bf16=False)
                flat[k] = old - h
                lm, _ = mlp.loss_and_grads(params, x, t, round_bf16=False)
                flat[k] = old
                assert abs((lp - lm) / (2 * h) - gf[k]) < 1e-7 * max(1.0, abs(gf[k]))
