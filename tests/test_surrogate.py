"""CPU checks of the branching-surrogate pieces: the numpy oracle of the
network (its backward pass against finite differences of its own loss, and
its BF16 rounding against hand-worked cases) and the candidate features."""
import numpy as np

from oracle import mlp_oracle as MO
from sovereign_amd import problems as P
from sovereign_amd import surrogate as S


def test_bf16_rounding_cases():
    # exactly representable values are unchanged
    v = np.array([1.0, -2.5, 0.0, 3.0 * 2 ** -10], np.float32)
    assert np.array_equal(MO.bf16(v), v.astype(np.float64))
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: even mantissa (1) wins
    assert MO.bf16(np.array([1.0 + 2 ** -8], np.float32))[0] == 1.0
    # 1 + 3 * 2^-8 ties between 1 + 2^-7 and 1 + 2^-6... rounds to even: 1 + 2^-6
    assert MO.bf16(np.array([1.0 + 3 * 2 ** -8], np.float32))[0] == 1.0 + 2 ** -6
    # just above the tie rounds up
    assert MO.bf16(np.array([1.0 + 2 ** -8 + 2 ** -20], np.float32))[0] == 1.0 + 2 ** -7
    # unit roundoff of an 8-bit significand: relative error <= 2^-8
    x = np.random.default_rng(0).normal(size=1000).astype(np.float32)
    assert np.max(np.abs(MO.bf16(x) - x) / np.abs(x)) <= 2 ** -8


def test_oracle_gradients_match_finite_differences():
    shape = (5, 7, 6)
    rng = np.random.default_rng(1)
    p = MO.he_init(shape, 2) + rng.normal(0, 0.05, MO.param_count(shape))
    X = rng.normal(size=(9, 5))
    t = rng.normal(size=9)
    loss, g = MO.loss_and_gradients(p, shape, X, t, round=False)
    eps = 1e-6
    num = np.zeros_like(p)
    for k in range(len(p)):
        e = np.zeros_like(p)
        e[k] = eps
        lp, _ = MO.loss_and_gradients(p + e, shape, X, t, round=False)
        lm, _ = MO.loss_and_gradients(p - e, shape, X, t, round=False)
        num[k] = (lp - lm) / (2 * eps)
    assert np.max(np.abs(num - g)) <= 1e-6 * max(1.0, np.max(np.abs(g)))
    # a planted mistake (ReLU mask dropped in layer 1) is caught by the same check
    W1, b1, W2, b2, w3, b3 = MO.unpack(p, shape)
    assert np.any(np.maximum(X @ W1 + b1, 0) != X @ W1 + b1)


def test_oracle_rounded_forward_close_to_exact():
    shape = (8, 16, 16)
    p = MO.he_init(shape, 3)
    X = np.random.default_rng(4).normal(size=(20, 8))
    _, _, y = MO.forward(p, shape, X, round=True)
    _, _, ye = MO.forward(p, shape, X, round=False)
    assert np.max(np.abs(y - ye)) <= 0.05 * max(1.0, np.max(np.abs(ye)))


def test_candidate_features_shape_and_ranges():
    model = P.multi_knapsack(40, 6, seed=3)
    stats = S.ModelStats(model)
    rng = np.random.default_rng(5)
    x = rng.uniform(0, 1, model.n_vars)
    cand = [0, 5, 17, 39]
    F = S.candidate_features(stats, x, model.lb, model.ub, cand)
    assert F.shape == (4, S.N_FEATURES)
    assert np.all(np.isfinite(F))
    frac = np.minimum(x - np.floor(x), np.ceil(x) - x)
    np.testing.assert_allclose(F[:, 0], frac[cand])
    assert np.all(F[:, 31] == 1.0)
    # most-fractional rank: 0 for the candidate closest to 1/2
    assert F[np.argmin(np.abs(frac[cand] - 0.5)), 30] == 0.0


def test_node_targets_and_inputs():
    cands = [{"score": 2.0}, {"score": 4.0}, {"score": 1e6}]
    t = S.node_targets(cands)
    assert t.max() == 1.0 and abs(t[0] - np.sqrt(2.0 / 1e6)) < 1e-15
    F = np.array([[1.0, 5.0, 2.0], [3.0, 5.0, 4.0], [5.0, 5.0, 0.0]])
    Z = S.node_inputs(F)
    np.testing.assert_allclose(Z.mean(axis=0), 0.0, atol=1e-12)
    assert np.all(Z[:, 1] == 0.0)  # node-wide feature carries nothing
    np.testing.assert_allclose(Z[:, 0], [-np.sqrt(1.5), 0.0, np.sqrt(1.5)])
