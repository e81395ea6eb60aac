"""Oracle pins for the structured-volume iso-surface (MCIsoSurfaceAccel over
StructuredDataSampler): closed forms the macrocell walk must reproduce.

* A field linear in world position is reproduced exactly by trilinear
  interpolation, and the linear root inside a sub-step is then exact: the
  iso-surface is a plane, hit at the analytic t, with the plane's normal.
* A radial field |p - c| gives a sphere of radius = iso value: t within the
  interpolation error of a voxel, normal radial.
* Values outside the field's range are never hit; macrocells whose range
  excludes the value are skipped without changing the first crossing."""
import numpy as np

from oracle import refcpu


def _rays(org, d, tmax=1e30):
    d = d / np.linalg.norm(d, axis=1, keepdims=True)
    n = len(org)
    return np.concatenate([org, np.zeros((n, 1)), d, np.full((n, 1), tmax)], 1).astype(np.float32)


def test_iso_plane_of_linear_fields():
    n = 24
    g = np.arange(n, dtype=np.float64)
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    rng = np.random.default_rng(3)
    m = 300
    for coef in ([1.0, 0.0, 0.0], [1.0, 1.0, 1.0], [0.3, -0.5, 0.8]):
        a = np.array(coef) / np.abs(coef).sum() / (n - 1)
        f = (a[0] * x + a[1] * y + a[2] * z).astype(np.float32)
        f -= f.min()
        v = refcpu.Volume(f, cells_per_mc=4)
        iso = float(f.max()) * 0.5
        # rays from outside the box toward random interior points
        tgt = rng.uniform(2, n - 3, (m, 3))
        org = tgt + rng.normal(size=(m, 3)) * 3 * n
        out = v.iso_rays(_rays(org, tgt - org), iso)
        d = (tgt - org) / np.linalg.norm(tgt - org, axis=1, keepdims=True)
        # s(p) = a.p - min; crossing where a.(o + t d) = iso + min_raw
        min_raw = (a * np.where(a < 0, n - 1, 0)).sum()
        t_plane = (iso + min_raw - org @ a) / (d @ a)
        # rays whose plane hit lies in the box, at least one sub-step (< 1
        # voxel here) inside the upper faces: the sub-step sample on a face
        # at rel = n - 1 is NaN (outside the cell grid) and cannot bracket
        # a crossing, as in the reference
        P = org + t_plane[:, None] * d
        inside = np.all((P > 0.01) & (P < n - 2), axis=1) & (t_plane > 0)
        assert inside.sum() > 50
        hit = np.isfinite(out[:, 0])
        assert hit[inside].all()
        np.testing.assert_allclose(out[inside, 0], t_plane[inside], rtol=1e-4, atol=1e-3)
        nrm = a / np.linalg.norm(a)
        np.testing.assert_allclose(out[inside, 1:], np.tile(nrm, (inside.sum(), 1)), atol=2e-3)
        # a value outside [min, max] is never hit
        assert not np.isfinite(v.iso_rays(_rays(org, tgt - org), float(f.max()) + 1.0)[:, 0]).any()


def test_iso_sphere_of_radial_field():
    n = 48
    c = np.array([23.5, 23.5, 23.5])
    g = np.arange(n, dtype=np.float64)
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    f = np.sqrt((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2).astype(np.float32)
    v = refcpu.Volume(f, cells_per_mc=8)
    R = 15.0
    rng = np.random.default_rng(5)
    m = 400
    u = rng.normal(size=(m, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    org = c + 60.0 * u
    tgt = c + rng.uniform(-8, 8, (m, 3))
    rays = _rays(org, tgt - org)
    out = v.iso_rays(rays, R)
    d = rays[:, 4:7].astype(np.float64)
    oc = org - c
    b = (oc * d).sum(1)
    disc = b * b - ((oc * oc).sum(1) - R * R)
    t_sphere = -b - np.sqrt(disc)
    assert (disc > 0).all() and np.isfinite(out[:, 0]).all()
    # trilinear interpolation of |p - c| at radius R errs by up to about
    # 3 h^2 / (8 R) = 0.025 voxel (h = 1): a 2x margin
    np.testing.assert_allclose(out[:, 0], t_sphere, atol=0.05)
    P = org + out[:, :1] * d
    radial = (P - c) / np.linalg.norm(P - c, axis=1, keepdims=True)
    assert np.all((out[:, 1:] * radial).sum(1) > 0.999)
