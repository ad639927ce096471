"""Pointwise probes of the device lattice primitives (the reference tests'
``moment_probe``/``phase_probe``/``constitutive_probe`` kernels,
iceflow3d/tests/test_home_phase.py:36-66, test_numerics.py:66-80)."""

import ctypes

import numpy as np
import torch

from .._native import check, lib


def _dev(array):
    return torch.from_numpy(np.ascontiguousarray(array, dtype=np.float32)).cuda()


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def home_probe(data):
    """data [n, 15]: m[10], acceleration[3], rate, bulk -> (populations[n,27], collided[n,10])."""
    d = _dev(data)
    rows = d.shape[0]
    pops = torch.empty((rows, 27), dtype=torch.float32, device=d.device)
    post = torch.empty((rows, 10), dtype=torch.float32, device=d.device)
    check(lib().icy3d_probe_home(_ptr(d), rows, _ptr(pops), _ptr(post), _stream()), "probe_home")
    return pops.cpu().numpy(), post.cpu().numpy()


def phase_probe(data, mobility, width):
    """data [n, 14]: h[7], phi, u[3], normal[3] -> [n, 10] (post[7], flux[3])."""
    d = _dev(data)
    rows = d.shape[0]
    out = torch.empty((rows, 10), dtype=torch.float32, device=d.device)
    check(lib().icy3d_probe_phase(_ptr(d), rows, float(mobility), float(width), _ptr(out), _stream()), "probe_phase")
    return out.cpu().numpy()


def viscosity_probe(phi, temps, rho_water, rho_air, mu_air):
    """Mixture kinematic viscosity (SI) on the phi x T grid."""
    p, t = _dev(phi), _dev(temps)
    out = torch.empty((p.shape[0], t.shape[0]), dtype=torch.float32, device=p.device)
    check(lib().icy3d_probe_viscosity(_ptr(p), p.shape[0], _ptr(t), t.shape[0], float(rho_water), float(rho_air),
                                      float(mu_air), _ptr(out), _stream()), "probe_viscosity")
    return out.cpu().numpy().astype(float)
