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
    """The GPU adjoint's gradient entries against central differences of the
    loss, 1e-3 relative: the 6 largest transfer-function entries and the 6
    largest density voxels. The differences are taken of the float64 march
    (same float32 sample geometry as the kernels) with h = 1e-5, where
    truncation and rounding are far below 1e-3 (a float32 forward's loss
    noise over h would not be)."""
    from barney_amd import ea
    dens, tf, cams = _scene(16, seed=1)
    W = H = 16
    n = dens.shape[0]
    cams = cams[:2]
    setup = ea.EASetup((n, n, n), cams, W, H, step=0.5, base_density=0.8)
    target = np.random.default_rng(3).uniform(0.0, 0.4, (2, H, W, 3))
    dg = torch.tensor(dens, device="cuda", requires_grad=True)
    tg = torch.tensor(tf, device="cuda", requires_grad=True)
    ((ea.ea_render(dg, tg, setup) - torch.tensor(target, device="cuda", dtype=torch.float32)) ** 2).sum().backward()
    plans = []
    for c in cams:
        cam = ea_torch.camera_dd(c["position"], c["direction"], c.get("up", (0, 1, 0)), c["fovy"])
        org, d = ea_torch.camera_rays(cam, W, H)
        plans.append(ea_torch.sample_plan((n, n, n), (0, 0, 0), (1, 1, 1), 0.5, org, d))
    tgt = torch.from_numpy(target)

    def loss(d, t):
        img = torch.stack([ea_torch.march(d, t, pl, W * H, 0.5, base_density=0.8).reshape(H, W, 3) for pl in plans])
        return float(((img - tgt) ** 2).sum())

    d64 = torch.tensor(dens, dtype=torch.float64)
    t64 = torch.tensor(tf, dtype=torch.float64)
    h = 1e-5
    checked = 0
    for name, x64, grad in (("tf", t64, tg.grad), ("density", d64, dg.grad)):
        g = grad.cpu().double().numpy().ravel()
        for k in np.argsort(-np.abs(g))[:6]:
            e = torch.zeros(x64.numel(), dtype=torch.float64)
            e[k] = h
            e = e.reshape(x64.shape)
            if name == "tf":
                fd = (loss(d64, t64 + e) - loss(d64, t64 - e)) / (2 * h)
            else:
                fd = (loss(d64 + e, t64) - loss(d64 - e, t64)) / (2 * h)
            print(f"{name}[{k}] gpu {g[k]:.6e} fd {fd:.6e} rel {abs(fd - g[k]) / abs(fd):.2e}", flush=True)
            assert abs(fd - g[k]) <= 1e-3 * abs(fd), (name, k, fd, g[k])
            checked += 1
    assert checked == 12
