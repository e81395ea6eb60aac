"""Shared test helpers: small scenes in both the product's and the oracle's
input form, and ray sets."""
import numpy as np

from barney_amd import scenes


def small_sphere_scene(num_spheres=48, stacks=8, slices=8, seed=0x5EED0000, extent=30.0):
    """Tessellated spheres (workload-1 style, reduced)."""
    c = scenes.uniform(seed, 1, 3 * num_spheres, 0.0, extent).reshape(-1, 3)
    r = scenes.uniform(seed, 2, num_spheres, 0.5, 2.0)
    col = scenes.uniform(seed, 3, 3 * num_spheres, 0.2, 0.9).reshape(-1, 3)
    out = []
    for s in range(num_spheres):
        v, t = scenes.sphere_mesh(c[s], r[s], stacks, slices)
        out.append(dict(kind=0, vertices=v, indices=t, color=col[s].astype(np.float32), user_id=s))
    return out


def user_prim_scene(n_spheres=200, n_cyl=200, seed=0x5EED00AA, extent=20.0):
    sc = scenes.uniform(seed, 1, 3 * n_spheres, 0.0, extent).reshape(-1, 3).astype(np.float32)
    sr = scenes.uniform(seed, 2, n_spheres, 0.2, 1.0).astype(np.float32)
    a = scenes.uniform(seed, 3, 3 * n_cyl, 0.0, extent).reshape(-1, 3)
    d = scenes.uniform(seed, 4, 3 * n_cyl, -2.0, 2.0).reshape(-1, 3)
    cv = np.concatenate([a, a + d]).astype(np.float32)
    ci = np.stack([np.arange(n_cyl), np.arange(n_cyl) + n_cyl], 1).astype(np.int32)
    cr = scenes.uniform(seed, 5, n_cyl, 0.05, 0.4).astype(np.float32)
    return [dict(kind=1, vertices=sc, radii=sr, color=(0.7, 0.3, 0.3), user_id=7),
            dict(kind=2, vertices=cv, indices=ci, radii=cr, color=(0.3, 0.7, 0.3), user_id=8)]


def to_product_geoms(geoms):
    from barney_amd import api
    out, mats = [], []
    for k, g in enumerate(geoms):
        mats.append(api.matte(g.get("color", (0.8, 0.8, 0.8))))
        out.append(api.Geom(g["kind"], g["vertices"], g.get("indices"), g.get("radii"), g.get("colors"),
                            g.get("default_radius", 1.0), material=k, user_id=g.get("user_id", 0)))
    return out, mats


def random_rays(n, lo, hi, seed=7, tmax=1e30):
    """Rays from random points on a sphere around the box toward random
    points inside it."""
    lo, hi = np.asarray(lo, np.float64), np.asarray(hi, np.float64)
    c = 0.5 * (lo + hi)
    rad = 1.5 * np.linalg.norm(hi - lo)
    u = scenes.uniform(seed, 1, 3 * n, -1.0, 1.0).reshape(-1, 3)
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    org = c + rad * u
    tgt = lo + (hi - lo) * scenes.uniform(seed, 2, 3 * n).reshape(-1, 3)
    d = tgt - org
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = org
    rays[:, 3] = 0.0
    rays[:, 4:7] = d
    rays[:, 7] = tmax
    return rays


def camera_rays(cam, width, height):
    """Primary rays of the oracle camera for every pixel (accumID 0)."""
    import ctypes as C
    from oracle import refcpu
    L = refcpu.load()
    p = refcpu.render_params(width, height)
    rays = np.zeros((height * width, 8), np.float32)
    org = np.zeros(3, np.float32)
    d = np.zeros(3, np.float32)
    seed = C.c_uint64(0)
    for iy in range(height):
        for ix in range(width):
            L.ref_camera_ray(C.byref(cam), C.byref(p), ix, iy, org.ctypes.data, d.ctypes.data, C.byref(seed))
            r = rays[iy * width + ix]
            r[0:3] = org
            r[4:7] = d
            r[7] = 1e30
    return rays
