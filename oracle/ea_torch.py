"""FP64 PyTorch oracle for the emission-absorption march and its gradients.

TEST INFRASTRUCTURE ONLY (tests/ and bench.py's CPU-baseline leg).

The march restates the product's definition (include/bnamd.h, EA section):
camera rays as barney's generateRays.cu:73-129 (pixel center, accumID 0) with
the zero-component fix of traceRays.dev.cu:42-44; samples at
t0 + (i + 1/2) dt for i < floor((t1 - t0) / dt) over the grid box; density
by trilinear interpolation with embree's texel convention
(rtcore/embree/Texture.cpp:326-409, StructuredData.h:104-115); color and
extinction from TransferFunction::DD::map (TransferFunction.h:67-86);
C = sum T_i a_i c_i. Sample positions, cells and weights are computed in
float32 exactly as the GPU does; everything that depends on the density or
the transfer function is float64 and differentiated by torch autograd.
"""
import math

import numpy as np
import torch

f32 = np.float32


def _nrm(v):
    v = np.asarray(v, f32)
    return (v * f32(f32(1.0) / np.sqrt(np.float32(np.dot(v, v))))).astype(f32)


def _dot(a, b):
    return f32(f32(f32(a[0] * b[0]) + f32(a[1] * b[1])) + f32(a[2] * b[2]))


def _cross(a, b):
    return np.array([f32(a[1] * b[2]) - f32(a[2] * b[1]), f32(a[2] * b[0]) - f32(a[0] * b[2]),
                     f32(a[0] * b[1]) - f32(a[1] * b[0])], f32)


def camera_dd(position, direction, up, fovy):
    """PerspectiveCamera::commit (barney/Camera.cu:81-100) in float32."""
    def nrm(v):
        v = np.asarray(v, f32)
        s = f32(f32(1.0) / np.sqrt(_dot(v, v)))
        return (v * s).astype(f32)
    d = nrm(nrm(direction))
    du = nrm(_cross(d, np.asarray(up, f32)))
    dv = nrm(_cross(du, d))
    scale = f32(1.0) / (f32(2.0) * f32(math.tan(f32(f32(f32(0.5) * f32(fovy)) * f32(math.pi)) / f32(180.0))))
    return dict(org=np.asarray(position, f32), d00=(d * f32(scale)).astype(f32), du=du, dv=dv)


def camera_rays(cam, W, H):
    """float32 (org, dir) for every pixel, accumID 0 (pixel centers)."""
    ix, iy = np.meshgrid(np.arange(W, dtype=f32), np.arange(H, dtype=f32))
    u = ((ix + f32(0.5)) / f32(W)).astype(f32)
    v = ((iy + f32(0.5)) / f32(H)).astype(f32)
    aspect = f32(f32(W) / f32(H))
    a = (f32(1.0) * aspect * (u - f32(0.5))).astype(f32)
    b = (f32(1.0) * (v - f32(0.5))).astype(f32)
    d = (cam["d00"][None, None, :] + a[..., None] * cam["du"][None, None, :]).astype(f32)
    d = (d + b[..., None] * cam["dv"][None, None, :]).astype(f32)
    dd = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]).astype(f32) + d[..., 2] * d[..., 2]
    dd = dd.astype(f32)
    inv = (f32(1.0) / np.sqrt(dd)).astype(f32)
    d = (d * inv[..., None]).astype(f32)
    d = np.where(d == 0, f32(1e-6), d).astype(f32)
    org = np.broadcast_to(cam["org"], d.shape).astype(f32)
    return org.reshape(-1, 3), d.reshape(-1, 3)


def _lerp_coords(N, tc):
    x = (tc - f32(0.5)).astype(f32)
    i0 = np.floor(x).astype(np.int64)  # x > 0 here, int() == floor
    f = (x - i0.astype(f32)).astype(f32)
    lo = x <= 0
    hi = x >= f32(N - 1)
    i0 = np.where(lo, 0, np.where(hi, N - 1, i0))
    i1 = np.where(lo, 0, np.where(hi, N - 1, i0 + 1))
    f = np.where(lo | hi, f32(0.0), f).astype(f32)
    return i0, i1, f


def sample_plan(dims, origin, spacing, step, org, dirs):
    """float32 sample geometry: per ray the number of samples and, per
    sample, the 8 corner indices and trilinear weights (valid samples only).
    Returns (ray_index, corner_idx[S,8], weights[S,8] f64, n_per_ray)."""
    nx, ny, nz = dims
    origin = np.asarray(origin, f32)
    spacing = np.asarray(spacing, f32)
    lo = origin
    hi = (origin + spacing * np.array([nx - 1, ny - 1, nz - 1], f32)).astype(f32)
    inv = (f32(1.0) / dirs).astype(f32)
    a = ((lo[None] - org) * inv).astype(f32)
    b = ((hi[None] - org) * inv).astype(f32)
    t0 = np.maximum(f32(0.0), np.minimum(a, b).max(1)).astype(f32)
    t1 = np.maximum(a, b).min(1).astype(f32)
    n = np.where(t0 <= t1, np.floor(((t1 - t0) / f32(step)).astype(f32)), 0).astype(np.int64)
    n = np.maximum(n, 0)
    ray_idx = np.repeat(np.arange(len(org)), n)
    i = (np.concatenate([np.arange(k) for k in n]) if n.sum() else np.zeros(0, np.int64)).astype(f32)
    t = (t0[ray_idx] + ((i + f32(0.5)) * f32(step)).astype(f32)).astype(f32)
    P = (org[ray_idx] + (t[:, None] * dirs[ray_idx]).astype(f32)).astype(f32)
    rel = ((P - origin[None]) * (f32(1.0) / spacing)[None]).astype(f32)
    ok = (rel >= 0).all(1) & (rel[:, 0] < f32(nx - 1)) & (rel[:, 1] < f32(ny - 1)) & (rel[:, 2] < f32(nz - 1))
    ray_idx, rel = ray_idx[ok], rel[ok]
    x0, x1, fx = _lerp_coords(nx, (rel[:, 0] + f32(0.5)).astype(f32))
    y0, y1, fy = _lerp_coords(ny, (rel[:, 1] + f32(0.5)).astype(f32))
    z0, z1, fz = _lerp_coords(nz, (rel[:, 2] + f32(0.5)).astype(f32))
    def lin(x, y, z):
        return x + nx * (y + ny * z)
    corners = np.stack([lin(x0, y0, z0), lin(x1, y0, z0), lin(x0, y1, z0), lin(x1, y1, z0),
                        lin(x0, y0, z1), lin(x1, y0, z1), lin(x0, y1, z1), lin(x1, y1, z1)], 1)
    fx, fy, fz = (f.astype(np.float64) for f in (fx, fy, fz))
    ax, ay, az = 1 - fx, 1 - fy, 1 - fz
    w = np.stack([ax * ay * az, fx * ay * az, ax * fy * az, fx * fy * az,
                  ax * ay * fz, fx * ay * fz, ax * fy * fz, fx * fy * fz], 1)
    return ray_idx, corners, w, n


def march(density, tf, plan, num_rays, step, domain=(0.0, 1.0), base_density=1.0):
    """Differentiable FP64 march. density: (nz, ny, nx) float64 tensor,
    tf: (N, 4) float64 tensor. Returns (num_rays, 3) colors."""
    ray_idx, corners, w, _ = plan
    dens = density.reshape(-1)
    ri = torch.from_numpy(ray_idx)
    rho = (dens[torch.from_numpy(corners)] * torch.from_numpy(w)).sum(1)
    N = tf.shape[0]
    span = domain[1] - domain[0]
    fr = (rho - domain[0]) / span
    f = fr.clamp(0.0, 1.0) * (N - 1)
    idx = f.detach().floor().long().clamp(0, N - 2)
    frac = f - idx.to(f.dtype)
    m = (1 - frac)[:, None] * tf[idx] + frac[:, None] * tf[idx + 1]
    sigma = m[:, 3] * base_density
    alpha = 1 - torch.exp(-sigma * step)
    # exclusive cumulative product of (1 - alpha) per ray
    one_minus = 1 - alpha
    logT = torch.log(one_minus.clamp_min(1e-300))
    # samples of one ray are contiguous and in order
    csum = torch.cumsum(logT, 0)
    starts = np.zeros(num_rays + 1, np.int64)
    np.add.at(starts, ray_idx + 1, 1)
    starts = np.cumsum(starts)
    first = torch.from_numpy(starts[ray_idx])
    excl = csum - logT
    base = torch.where(first > 0, csum[(first - 1).clamp_min(0)], torch.zeros_like(csum))
    T = torch.exp(excl - base)
    contrib = (T * alpha)[:, None] * m[:, :3]
    out = torch.zeros(num_rays, 3, dtype=density.dtype)
    out = out.index_add(0, ri, contrib)
    return out


def render(density_np, tf_np, cams, W, H, step, domain=(0.0, 1.0), base_density=1.0, origin=(0, 0, 0),
           spacing=(1, 1, 1), requires_grad=False):
    """Renders all views; returns (images (V,H,W,3) tensor, density, tf)."""
    nz, ny, nx = density_np.shape
    density = torch.tensor(density_np, dtype=torch.float64, requires_grad=requires_grad)
    tf = torch.tensor(tf_np, dtype=torch.float64, requires_grad=requires_grad)
    imgs = []
    for c in cams:
        cam = camera_dd(c["position"], c["direction"], c.get("up", (0, 1, 0)), c["fovy"])
        org, d = camera_rays(cam, W, H)
        plan = sample_plan((nx, ny, nz), origin, spacing, step, org, d)
        imgs.append(march(density, tf, plan, W * H, step, domain, base_density).reshape(H, W, 3))
    return torch.stack(imgs), density, tf
