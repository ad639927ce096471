"""The learned closure end to end on the GPU: paired coarse / 2x-fine solver
runs make the data, the project's kernels train the network, and on a
held-out pair the trained network beats the plain coarse exchange, both on
the training objective (the smoothed fine/coarse heat-flux ratio) and on the
melted ice mass against the fine run."""

import pytest

pytestmark = pytest.mark.gpu

TRAIN_SEEDS = (11, 12, 13)
HELD_OUT = 21


def test_training_on_solver_pairs_beats_the_plain_exchange():
    from projecticy_amd.closure.network import ClosureNet
    from projecticy_amd.closure.train import baseline_loss, generate, mean_loss, melt_comparison, train

    samples, histories = generate(TRAIN_SEEDS, windows=3, window_steps=16)
    held, held_history = generate((HELD_OUT,), windows=3, window_steps=16)
    # the cubes start submerged: both resolutions melt ice in every window
    for history in list(histories.values()) + list(held_history.values()):
        assert history[-1][1] < history[0][1] and history[-1][2] < history[0][2]
    net = ClosureNet((64, 32, 32), seed=0)
    before = mean_loss(net, held)
    curve = train(net, samples, epochs=30, lr=2e-3)
    after = mean_loss(net, held)
    plain = baseline_loss(held)
    print(f"held-out loss: plain {plain:.4e}, initial net {before:.4e}, trained {after:.4e}; "
          f"training curve {curve[0]:.4e} -> {curve[-1]:.4e}")
    assert curve[-1] < curve[0]
    assert after < 0.8 * plain
    scores = melt_comparison(net, HELD_OUT, windows=3, window_steps=16)
    print("melted mass after 3 windows:", scores)
    assert scores["melted_fine_kg"] > 0.0 and scores["melted_plain_kg"] > 0.0
    # the coarse exchange under-resolves the thermal boundary layer; the learned gain closes most of the gap
    assert scores["closure_error"] < 0.5 * scores["plain_error"]
