"""The oracle of the step-proposal network, pinned without the product:
its analytic gradient against central finite differences (exact FP64 mode),
a forward pass computable by hand, and the bfloat16 rounding it emulates."""
import numpy as np

import oracle_ffi as O

NI, NH, NO = 24, 64, 2
OW1, OB1 = 0, NI * NH
OW2 = OB1 + NH
OB2 = OW2 + NH * NH
OW3 = OB2 + NH
OB3 = OW3 + NH * NO


def rand_problem(n=6, seed=0):
    rng = np.random.default_rng(seed)
    th = np.zeros(O.sn_param_count(), np.float32)
    th[OW1:OB1] = rng.normal(0, 0.3, NI * NH)
    th[OB1:OW2] = rng.normal(0, 0.1, NH)
    th[OW2:OB2] = rng.normal(0, 0.2, NH * NH)
    th[OB2:OW3] = rng.normal(0, 0.1, NH)
    th[OW3:OB3] = rng.normal(0, 0.2, NH * NO)
    th[OB3:] = rng.normal(0, 0.1, NO)
    x = rng.uniform(-1, 1, (n, NI)).astype(np.float32)
    t = np.stack([rng.uniform(0, 1, n), rng.integers(0, 2, n), rng.uniform(0.5, 2, n)], 1).astype(np.float32)
    return th, x, t


def test_param_count():
    assert O.sn_param_count() == 24 * 64 + 64 + 64 * 64 + 64 + 64 * 2 + 2 == 5890


def test_gradient_matches_finite_differences():
    th, x, t = rand_problem()
    g, _ = O.sn_grad(th, x, t, bf16=False)
    rng = np.random.default_rng(1)
    # every block of the parameter vector, a few entries each
    idx = np.concatenate([rng.choice(np.arange(a, b), 12, replace=False) if b - a > 12 else np.arange(a, b)
                          for a, b in [(OW1, OB1), (OB1, OW2), (OW2, OB2), (OB2, OW3), (OW3, OB3), (OB3, OB3 + 2)]])
    n = x.shape[0]
    for i in idx:
        h = np.float32(1e-3)
        tp, tm = th.copy(), th.copy()
        tp[i] += h
        tm[i] -= h
        lp = O.sn_grad(tp, x, t, bf16=False)[1] / n
        lm = O.sn_grad(tm, x, t, bf16=False)[1] / n
        fd = (lp - lm) / (np.float64(tp[i]) - np.float64(tm[i]))
        assert abs(fd - g[i]) <= 1e-5 + 1e-3 * abs(g[i]), (i, fd, g[i])


def test_forward_by_hand():
    th = np.zeros(O.sn_param_count(), np.float32)
    th[OW1 + 0 * NH + 0] = 1.0     # h1[0] = x[0]
    th[OW2 + 0 * NH + 5] = 2.0     # h2[5] = 2 h1[0]
    th[OB2 + 6] = -1.0             # h2[6] = relu(-1) = 0
    th[OW3 + 5 * NO + 0] = 0.5     # y0 = 0.5 h2[5]
    th[OW3 + 6 * NO + 1] = 7.0     # y1 = 7 h2[6] + b3[1]
    th[OB3 + 1] = 0.25
    x = np.zeros((2, NI), np.float32)
    x[0, 0], x[1, 0] = 0.75, -3.0
    for bf in (False, True):
        y = O.sn_forward(th, x, bf16=bf)
        assert np.array_equal(y, np.array([[0.75, 0.25], [0.0, 0.25]]))


def test_bf16_rounding_is_nearest_even():
    th = np.zeros(O.sn_param_count(), np.float32)
    th[OW1] = 1.0
    th[OW2] = 1.0
    th[OW3] = 1.0
    x = np.zeros((4, NI), np.float32)
    # 1 + 2^-8 is half way between bfloat16 neighbours 1 and 1 + 2^-7: ties go to the even (1)
    x[:, 0] = [1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -7, 1 + 2 ** -9]
    y = O.sn_forward(th, x, bf16=True)[:, 0]
    assert list(y) == [1.0, 1 + 2 ** -6, 1 + 2 ** -7, 1.0]
    assert list(O.sn_forward(th, x, bf16=False)[:, 0]) == list(x[:, 0].astype(np.float64))
