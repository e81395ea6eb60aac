"""GPU emission-absorption forward + checkpointed adjoint against the FP64
PyTorch oracle (forward 1e-4 relative; gradients 1e-3 relative in norm) and
against central finite differences of the GPU forward itself."""
import math

import numpy as np
import pytest

from oracle import ea_torch

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _scene(n=20, seed=0):
    from barney_amd import scenes
    dens = scenes.structured_field(n=n, blobs=5, sigma=(3.0, 8.0), octaves=1, seed=0x5EED0300 + seed)
    rng = np.random.default_rng(seed)
    x = np.linspace(0, 1, 32)
    tf = np.stack([0.5 + 0.5 * np.sin(3 * x), x, 1 - x, 0.8 * x ** 1.5], -1).astype(np.float32)
    tf[:, :3] += rng.uniform(-0.05, 0.05, (32, 3)).astype(np.float32)
    c = (n - 1) / 2
    cams = []
    for k in range(3):
        a = 2 * math.pi * (k + 0.1) / 3
        pos = (c + 2.2 * n * math.cos(a), c + 0.5 * n, c + 2.2 * n * math.sin(a))
        cams.append(dict(position=pos, direction=tuple(c - p for p in pos), fovy=30.0))
    return dens, tf, cams


@pytest.mark.parametrize("ckpt", [8, 16])
def test_forward_and_adjoint_vs_fp64(gpu, ckpt):
    from barney_amd import ea
    dens, tf, cams = _scene()
    W = H = 24
    n = dens.shape[0]
    setup = ea.EASetup((n, n, n), cams, W, H, step=0.5, base_density=0.8, checkpoint_every=ckpt)
    dg = torch.tensor(dens, device="cuda", requires_grad=True)
    tg = torch.tensor(tf, device="cuda", requires_grad=True)
    img = ea.ea_render(dg, tg, setup)
    ref, dr, tr = ea_torch.render(dens, tf, cams, W, H, 0.5, base_density=0.8, requires_grad=True)
    np.testing.assert_allclose(img.detach().cpu().numpy(), ref.detach().numpy(), rtol=1e-4, atol=1e-6)
    target = torch.from_numpy(np.random.default_rng(5).uniform(0, 0.5, ref.shape))
    ((ref - target) ** 2).sum().backward()
    ((img - target.float().cuda()) ** 2).sum().backward()
    for g, r, name in ((dg.grad, dr.grad, "density"), (tg.grad, tr.grad, "tf")):
        g = g.cpu().double().numpy()
        r = r.numpy()
        rel = np.linalg.norm(g - r) / np.linalg.norm(r)
        assert rel < 1e-3, (name, rel)
    assert setup.last_backward["samples"] == setup.last_forward["samples"]


def test_adjoint_vs_central_differences(gpu):
    from barney_amd import ea
    dens, tf, cams = _scene(16, seed=1)
    W = H = 16
    n = dens.shape[0]
    setup = ea.EASetup((n, n, n), cams[:2], W, H, step=0.5, base_density=0.8)
    target = torch.rand((2, H, W, 3), device="cuda", generator=torch.Generator("cuda").manual_seed(3)) * 0.4
    dg = torch.tensor(dens, device="cuda", requires_grad=True)
    tg = torch.tensor(tf, device="cuda", requires_grad=True)
    ((ea.ea_render(dg, tg, setup) - target) ** 2).sum().backward()

    def loss(d, t):
        with torch.no_grad():
            return float(((ea.ea_render(d, t, setup) - target) ** 2).sum().double())

    h = 2e-3
    for k in np.argsort(-np.abs(tg.grad.cpu().numpy()).ravel())[:6]:
        e = torch.zeros(tg.numel(), device="cuda")
        e[k] = h
        e = e.reshape(tg.shape)
        fd = (loss(dg.detach(), tg.detach() + e) - loss(dg.detach(), tg.detach() - e)) / (2 * h)
        g = float(tg.grad.reshape(-1)[k])
        assert abs(fd - g) <= 1e-2 * abs(g) + 1e-4, (k, fd, g)
