"""GPU checks of the branching surrogate: the BF16-MFMA forward pass and
gradients against the numpy oracle (same rounding points), training on the
GPU lowering the loss, and what the committed trained network gains over
the most-fractional order on held-out knapsack instances: the candidate
that full strong branching picks lands in the top 8 more often, and branch
and bound with 8 strong-branching candidates per node reaches the same
optimum (checked against scipy) with fewer nodes."""
import os

import numpy as np
import pytest

from oracle import mlp_oracle as MO
from sovereign_amd import milp as MI
from sovereign_amd import problems as P
from sovereign_amd import surrogate as S

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[case] {msg}", flush=True)


@pytest.fixture(scope="module")
def net(gpu_api):
    n = S.Mlp(S.SHAPE, seed=11)
    yield n
    n.close()


def test_forward_matches_oracle(net):
    say("mlp forward vs oracle")
    rng = np.random.default_rng(1)
    X = rng.normal(size=(300, S.N_FEATURES)).astype(np.float32)
    y = net.forward(X)
    _, _, yo = MO.forward(net.params.astype(np.float64), S.SHAPE, X, round=True)
    assert np.max(np.abs(y - yo)) <= 1e-4 * max(1.0, np.max(np.abs(yo)))


def test_gradients_match_oracle(net):
    say("mlp gradients vs oracle")
    rng = np.random.default_rng(2)
    X = rng.normal(size=(257, S.N_FEATURES)).astype(np.float32)  # a partial 64-row tile
    t = rng.normal(size=257).astype(np.float32)
    loss, g = net.gradients(X, t)
    lo, go = MO.loss_and_gradients(net.params.astype(np.float64), S.SHAPE, X, t, round=True)
    assert abs(loss - lo) <= 1e-5 * max(1.0, lo)
    W = MO.unpack(g.astype(np.float64), S.SHAPE)
    Wo = MO.unpack(go, S.SHAPE)
    for a, b in zip(W, Wo):  # each block to 1e-3 of its own scale
        assert np.max(np.abs(a - b)) <= 1e-3 * max(np.max(np.abs(b)), 1e-6)
    # and the rounded gradients stay close to the exact FP64 ones
    _, ge = MO.loss_and_gradients(net.params.astype(np.float64), S.SHAPE, X, t, round=False)
    assert np.linalg.norm(g - ge) <= 0.05 * np.linalg.norm(ge)


def test_training_lowers_loss(gpu_api):
    say("mlp train on a planted target")
    rng = np.random.default_rng(3)
    X = rng.normal(size=(2048, S.N_FEATURES)).astype(np.float32)
    t = (np.tanh(X[:, 0] - 0.5 * X[:, 3]) + 0.3 * X[:, 7] ** 2).astype(np.float32)
    n = S.Mlp(S.SHAPE, seed=5)
    y0 = n.forward(X)
    hist = n.train(X, t, epochs=30, batch=256, lr=1e-3)
    y1 = n.forward(X)
    n.close()
    assert hist[-1] < 0.1 * hist[0]
    assert np.mean((y1 - t) ** 2) < 0.1 * np.mean((y0 - t) ** 2)


def test_trained_surrogate_gains_over_most_fractional(gpu_api):
    """On the 200 held-out instances recorded with the weights, the candidate
    full strong branching picks is in the network's top 8 more often than in
    the most-fractional top 8 (deterministic: same seeds, same solves)."""
    if not os.path.exists(S.WEIGHTS):
        pytest.skip("no trained weights committed")
    z = np.load(S.WEIGHTS, allow_pickle=False)
    seeds = [int(s) for s in z["eval_seeds"]]
    net = S.Mlp.load(S.WEIGHTS)
    ipm = S.InteriorPointSolver(S.InteriorPointOptions(feasibility_tol=1e-9, optimality_tol=1e-9))
    hits_s = hits_f = nodes = 0
    say(f"surrogate ranking on {len(seeds)} held-out instances")
    for seed in seeds:
        model = S.instance(seed)
        root = ipm.solve(model)
        F, t, _ = S.node_dataset(model, S.ModelStats(model), root["primal"], root["objective"])
        if len(t) < 10:
            continue
        nodes += 1
        best = int(np.argmax(t))
        ys = net.forward(S.node_inputs(F))
        hits_s += best in np.argsort(-ys, kind="stable")[:8]
        hits_f += best in np.argsort(np.abs(F[:, 0] - 0.5), kind="stable")[:8]
    net.close()
    say(f"top-8 recall of the strong-branching choice: surrogate {hits_s}/{nodes}, most fractional {hits_f}/{nodes}")
    assert hits_s > hits_f


def test_branch_and_bound_with_surrogate_same_optimum(gpu_api):
    """Branch and bound with 8 strong-branching candidates per node, ranked by
    the trained network or most-fractional first, and with full strong
    branching: all reach scipy's optimum."""
    if not os.path.exists(S.WEIGHTS):
        pytest.skip("no trained weights committed")
    from scipy.optimize import Bounds, LinearConstraint, milp as scipy_milp

    model = P.multi_knapsack(30, 6, seed=77)
    is_int = np.ones(model.n_vars, dtype=np.int8)
    r = scipy_milp(-model.obj, constraints=LinearConstraint(model.dense(), -np.inf, model.rhs),
                   integrality=np.ones(model.n_vars), bounds=Bounds(model.lb, model.ub))
    opt = -r.fun
    ranker = S.load_trained(model)
    runs = {}
    for name, kw in (("full", dict(strong_branch_candidates=64)),
                     ("most-fractional top-8", dict(strong_branch_candidates=8)),
                     ("surrogate top-8", dict(strong_branch_candidates=8, ranker=ranker))):
        say(f"bnb mkp 30x6 {name}")
        runs[name] = MI.BranchAndBoundSolver(branch_rule="full_strong", **kw).solve(model, is_int)
    ranker.net.close()
    for name, res in runs.items():
        say(f"{name}: nodes {res['nodes']} lp solves {res['lp_solves']} objective {res['objective']}")
        assert res["status"] == 0
        assert abs(res["objective"] - opt) <= 1e-6 * max(1.0, abs(opt))
    assert ranker.calls > 0
