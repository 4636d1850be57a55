"""The warm-start network's CPU oracle pinned by mathematics: its analytic gradients against central
differences of its own loss (plain f64 mode), its BF16 rounding against exact cases, and its Adam step
against the closed form of the first step (|update| = lr for every nonzero gradient)."""
import numpy as np

from oracle import mlp_oracle as M


def _net(widths, seed=0, scale=0.5):
    rng = np.random.default_rng(seed)
    return [(rng.normal(0, scale, (widths[l + 1], widths[l])), rng.normal(0, 0.1, widths[l + 1]))
            for l in range(len(widths) - 1)]


def test_gradients_match_central_differences():
    widths = [3, 7, 6, 5, 4]
    params = _net(widths)
    rng = np.random.default_rng(1)
    x = rng.uniform(-1, 1, (5, 3))
    t = rng.normal(0, 1, (5, 4))
    base = rng.normal(0, 0.1, 4)
    sc = 0.3
    loss, grads = M.loss_and_grads(params, x, t, base, bf=False, scale=sc)
    h = 1e-6
    for l, (w, b) in enumerate(params):
        for arr, g in ((w, grads[l][0]), (b, grads[l][1])):
            num = np.zeros_like(arr)
            it = np.nditer(arr, flags=["multi_index"])
            for _ in it:
                i = it.multi_index
                old = arr[i]
                arr[i] = old + h
                lp, _ = M.loss_and_grads(params, x, t, base, bf=False, scale=sc)
                arr[i] = old - h
                lm, _ = M.loss_and_grads(params, x, t, base, bf=False, scale=sc)
                arr[i] = old
                num[i] = (lp - lm) / (2 * h)
            assert np.max(np.abs(num - g)) < 1e-7 * max(1.0, np.max(np.abs(g))), (l, np.max(np.abs(num - g)))
    # the loss is the per-sample squared error in units of the output scale
    z = M.forward(params, x, bf=False)[-1]
    assert abs(loss - np.sum(((base + sc * z - t) / sc) ** 2) / len(x)) < 1e-12


def test_bf16_rounding():
    # exactly representable values stay; ties go to even; 1 + 2^-8 rounds to 1, 1 + 3 * 2^-9 to 1 + 2^-7
    v = np.array([1.0, -2.5, 0.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -9, 1.0 + 2.0 ** -7 + 2.0 ** -8, 3.140625])
    r = M.bf16(v)
    assert np.array_equal(r, [1.0, -2.5, 0.0, 1.0, 1.0 + 2.0 ** -7, 1.0 + 2.0 ** -6, 3.140625])


def test_adam_first_step_is_lr_sign():
    p = np.array([1.0, -2.0, 3.0])
    g = np.array([0.3, -1e-3, 0.0])
    p1, m, v = M.adam_step(p, g, np.zeros(3), np.zeros(3), 1, lr=0.01, eps=0.0 + 1e-30)
    assert np.allclose(p1[:2] - p[:2], -0.01 * np.sign(g[:2]), rtol=1e-12)
    assert p1[2] == p[2]
