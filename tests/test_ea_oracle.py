"""Pins the FP64 emission-absorption oracle: closed-form transmittance for a
homogeneous grid, and its autograd gradients against central finite
differences (both in float64, on the CPU)."""
import math

import numpy as np
import pytest
import torch

from oracle import ea_torch


def _cams(n=2, r=30.0, c=(5.5, 5.5, 5.5)):
    out = []
    for k in range(n):
        a = 2 * math.pi * (k + 0.3) / n
        pos = (c[0] + r * math.cos(a), c[1] + 0.4 * r, c[2] + r * math.sin(a))
        out.append(dict(position=pos, direction=tuple(c[i] - pos[i] for i in range(3)), fovy=35.0))
    return out


def test_homogeneous_closed_form():
    # constant density rho -> every sample has the same (c, sigma):
    # C = c * (1 - exp(-sigma dt)^n) with n = floor((t1 - t0) / dt)
    nx = 12
    dens = np.full((nx, nx, nx), 0.5, np.float32)
    tf = np.array([[0.2, 0.4, 0.6, 0.1], [0.2, 0.4, 0.6, 0.3]], np.float32)  # at 0.5: w = 0.2
    cam = dict(position=(5.5, 5.5, -20.0), direction=(0, 0, 1), fovy=10.0)
    step = 0.5
    img, _, _ = ea_torch.render(dens, tf, [cam], 4, 4, step, base_density=2.0)
    camd = ea_torch.camera_dd(cam["position"], cam["direction"], (0, 1, 0), cam["fovy"])
    org, d = ea_torch.camera_rays(camd, 4, 4)
    plan = ea_torch.sample_plan((nx, nx, nx), (0, 0, 0), (1, 1, 1), step, org, d)
    n = plan[3]
    sigma = 0.2 * 2.0
    for k in range(16):
        ek = 1 - math.exp(-sigma * step) ** n[k]
        # tolerance covers the float32 rounding of the TF entries
        np.testing.assert_allclose(img[0].reshape(-1, 3)[k].numpy(), np.array([0.2, 0.4, 0.6]) * ek, rtol=1e-6)
    assert n.min() > 10


def test_oracle_gradients_match_central_differences():
    rng = np.random.default_rng(0)
    n = 10
    dens = rng.uniform(0.0, 1.0, (n, n, n)).astype(np.float32)
    tf = np.concatenate([rng.uniform(0, 1, (16, 3)), rng.uniform(0, 0.6, (16, 1))], 1).astype(np.float32)
    target = torch.from_numpy(rng.uniform(0, 0.3, (2, 6, 6, 3)))
    cams = _cams(2, r=25.0, c=(4.5, 4.5, 4.5))

    def loss(d, t):
        img, _, _ = ea_torch.render(d, t, cams, 6, 6, 0.5, base_density=1.5)
        return ((img - target) ** 2).sum()

    img, dt, tt = ea_torch.render(dens, tf, cams, 6, 6, 0.5, base_density=1.5, requires_grad=True)
    L = ((img - target) ** 2).sum()
    L.backward()
    gd, gt = dt.grad.numpy(), tt.grad.numpy()
    # central differences on a few coordinates (float64 perturbations of the
    # float64 parameters: render() takes float32 inputs, so perturb through
    # the tensors directly)
    def loss_t(d64, t64):
        d64 = d64.detach().clone()
        t64 = t64.detach().clone()
        img = []
        for c in cams:
            cam = ea_torch.camera_dd(c["position"], c["direction"], (0, 1, 0), c["fovy"])
            org, dd = ea_torch.camera_rays(cam, 6, 6)
            plan = ea_torch.sample_plan((n, n, n), (0, 0, 0), (1, 1, 1), 0.5, org, dd)
            img.append(ea_torch.march(d64, t64, plan, 36, 0.5, base_density=1.5).reshape(6, 6, 3))
        return float(((torch.stack(img) - target) ** 2).sum())

    d64 = torch.tensor(dens, dtype=torch.float64)
    t64 = torch.tensor(tf, dtype=torch.float64)
    h = 1e-6
    idx = np.argsort(-np.abs(gd).ravel())[:6]
    for i in idx:
        e = torch.zeros(d64.numel(), dtype=torch.float64)
        e[i] = h
        fd = (loss_t(d64 + e.reshape(d64.shape), t64) - loss_t(d64 - e.reshape(d64.shape), t64)) / (2 * h)
        assert abs(fd - gd.ravel()[i]) <= 1e-5 * max(1.0, abs(fd)), (i, fd, gd.ravel()[i])
    for i in np.argsort(-np.abs(gt).ravel())[:6]:
        e = torch.zeros(t64.numel(), dtype=torch.float64)
        e[i] = h
        fd = (loss_t(d64, t64 + e.reshape(t64.shape)) - loss_t(d64, t64 - e.reshape(t64.shape))) / (2 * h)
        assert abs(fd - gt.ravel()[i]) <= 1e-5 * max(1.0, abs(fd)), (i, fd, gt.ravel()[i])
