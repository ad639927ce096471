"""Learned interface closure for the 3D heat exchange (no reference counterpart).

A four-layer 3x3x3 CNN (32 channels, BF16 weights and activations, FP32
accumulation on the MI355X matrix cores) reads the coarse solver's phase,
water temperature, velocity and ice coverage and predicts a per-cell gain on
the ice/water heat exchange of iceflow3d/thermal.py:308-392.  It is trained
on paired coarse / 2x-fine runs of the solver itself, with the project's own
forward, backward, loss and Adam kernels (``libicyclosure.so``).  Without a
network the coarse solver runs the plain exchange (gain 1 everywhere).
"""

from .network import ClosureNet

__all__ = ["ClosureNet"]
