"""Generate paired data, train the closure network and score it.

Every step runs on the GPU: the solver makes the data, the project's
kernels run the network's forward and backward passes and Adam."""

import numpy as np
import torch

from .data import ClosureCoupling, closure_pair, ice_mass_kg, record_pair
from .network import ClosureNet


def generate(seeds, *, windows=4, window_steps=16, **pair_options):
    """Samples of every seeded pair, and each pair's melt history."""
    samples, histories = [], {}
    for seed in seeds:
        coarse, fine = closure_pair(seed, **pair_options)
        pair_samples, history = record_pair(coarse, fine, windows=windows, window_steps=window_steps)
        samples.extend(pair_samples)
        histories[seed] = history
        torch.cuda.empty_cache()
    return samples, histories


def baseline_loss(samples):
    """Loss of the plain method (prediction 0, i.e. gain 1) on the samples."""
    return float(np.mean([float((w.double() * t.double() ** 2).sum().item()) for _, t, w in samples]))


def mean_loss(net, samples):
    return float(np.mean([net.evaluate(f, t, w) for f, t, w in samples]))


def train(net, samples, *, epochs=40, lr=1e-3, seed=0):
    """Adam over shuffled samples; returns the mean loss of every epoch."""
    rng = np.random.default_rng(seed)
    history = []
    for _ in range(int(epochs)):
        losses = []
        for k in rng.permutation(len(samples)):
            losses.append(net.train_step(*samples[k], lr=lr))
        history.append(float(torch.stack(losses).mean().item()))
    return history


def melt_comparison(net, seed, *, windows=4, window_steps=16, **pair_options):
    """Melted ice mass after ``windows`` windows: the fine run, the plain
    coarse run and the coarse run with the network's gain."""
    from ..iceflow3d import IceFlow3D

    coarse_cfg, fine_cfg = closure_pair(seed, **pair_options)
    fine = IceFlow3D(fine_cfg)
    start_f = ice_mass_kg(fine)
    fine.step(2 * windows * window_steps)
    melted_fine = start_f - ice_mass_kg(fine)
    del fine
    plain = IceFlow3D(coarse_cfg)
    start_c = ice_mass_kg(plain)
    plain.step(windows * window_steps)
    melted_plain = start_c - ice_mass_kg(plain)
    del plain
    learned = IceFlow3D(coarse_cfg)
    coupling = ClosureCoupling(learned, net, window_steps=window_steps)
    coupling.run(windows)
    melted_learned = start_c - ice_mass_kg(learned)
    coupling.detach()
    scale = max(abs(melted_fine), 1e-30)
    return {
        "melted_fine_kg": melted_fine, "melted_plain_kg": melted_plain, "melted_closure_kg": melted_learned,
        "plain_error": abs(melted_plain - melted_fine) / scale,
        "closure_error": abs(melted_learned - melted_fine) / scale,
    }


def make_net(shape, seed=0):
    return ClosureNet(shape, seed=seed)
