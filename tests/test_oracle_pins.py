"""Pins the CPU oracle against what the mathematics fixes (no golden vectors
exist in the reference for this path): closed forms, textbook routines
computed a different way, brute force and invariants. Each test is chosen so
that a plausible slip in the oracle (a dropped term, a swapped operand, a
wrong constant) fails it."""
import math

import numpy as np
import pytest

from oracle import refcpu
from helpers import small_sphere_scene, user_prim_scene, random_rays

L = refcpu.load()
MASK64 = (1 << 64) - 1
A = 6364136223846793005


# ---------------------------------------------------------------------------
# RNG: barney's RNGSeed is a multiplicative LCG (Knuth's MMIX multiplier) with
# PCG's XSH-RS output (O'Neill 2014, "pcg32_fast"): out = (x ^ (x >> 22)) >>
# (22 + (x >> 61)) of the pre-advance state.
# ---------------------------------------------------------------------------

def test_rng_state_closed_form_and_output():
    import ctypes as C
    s0 = 0x0123456789ABCDEF
    st = C.c_uint64(s0)
    out = np.zeros(1000, np.uint32)
    L.ref_rng_sequence(C.byref(st), out.ctypes.data, 1000)
    assert st.value == (s0 * pow(A, 1000, 1 << 64)) & MASK64
    for n in (0, 1, 17, 999):
        x = (s0 * pow(A, n, 1 << 64)) & MASK64
        assert int(out[n]) == ((x ^ (x >> 22)) >> (22 + (x >> 61))) & 0xFFFFFFFF


def test_rng_seed_and_uniformity():
    # Random2(a, b): state ^= 2a+1, advance, state ^= 2b, advance
    s = L.ref_rng_seed(0xcafef00dd15ea5e5, 5, 9)
    x = 0xcafef00dd15ea5e5 ^ 11
    x = (x * A) & MASK64
    x ^= 18
    x = (x * A) & MASK64
    assert s == x
    f = np.zeros(200000, np.float32)
    L.ref_random2_floats(3, 4, f.ctypes.data, len(f))
    assert f.min() >= 0.0 and f.max() < 1.0
    assert np.all(np.floor(f * (1 << 24)) == f * (1 << 24))  # 24-bit grid
    hist, _ = np.histogram(f, bins=64, range=(0, 1))
    chi2 = ((hist - len(f) / 64) ** 2 / (len(f) / 64)).sum()
    assert chi2 < 120  # 63 dof, p ~ 1e-5
    assert L.ref_hash3(1, 2, 3) == ((((0xcbf29ce484222325 ^ (0x100000001b3 * 1)) * 0x100000001b3 ^ 2)
                                    * 0x100000001b3 ^ 3) & MASK64)


def test_fast_log_series():
    # 2*atanh((f-1)/(f+1)) truncated after the z^5 term: exact for f = 1,
    # error grows as f moves away from 1, always below |log f|
    assert L.ref_fast_log(1.0) == 0.0
    for f in (0.999, 0.9, 0.5):
        z = (f - 1) / (f + 1)
        series = 2 * (z + z ** 3 / 3 + z ** 5 / 5)
        assert abs(L.ref_fast_log(f) - series) < 1e-6
        assert abs(L.ref_fast_log(f)) <= abs(math.log(f)) + 1e-7
    assert abs(L.ref_fast_log(0.9) - math.log(0.9)) < 1e-6
    assert abs(L.ref_fast_log(0.01) - math.log(0.01)) > 1.0  # large bias for small (1-r)


# ---------------------------------------------------------------------------
# Intersectors vs double-precision references computed differently
# ---------------------------------------------------------------------------

def test_triangle_vs_f64_moller_trumbore():
    rng = np.random.default_rng(0)
    n = 3000
    tris = rng.uniform(-1, 1, (n, 3, 3)).astype(np.float32)
    geoms = [dict(kind=0, vertices=tris.reshape(-1, 3), indices=np.arange(3 * n, dtype=np.int32).reshape(-1, 3))]
    s = refcpu.Scene(geoms)
    org = rng.uniform(-3, 3, (n, 3)).astype(np.float32)
    tgt = (tris.mean(1) + rng.uniform(-0.3, 0.3, (n, 3))).astype(np.float32)
    checked = 0
    for i in range(n):
        d = tgt[i] - org[i]
        d = (d / np.linalg.norm(d)).astype(np.float32)
        ray = np.array([*org[i], 0, *d, 1e30], np.float32)
        # only this triangle: build per-triangle expectations by brute force
        args = [np.ascontiguousarray(x, np.float64) for x in (org[i], d, tris[i, 0], tris[i, 1], tris[i, 2])]
        t64 = L.ref_triangle_t_f64(*(a.ctypes.data for a in args))
        single = refcpu.Scene([dict(kind=0, vertices=tris[i], indices=np.array([[0, 1, 2]], np.int32))])
        h = single.trace(ray[None])
        if t64 > 1e-3:
            # skip rays within 1e-4 (barycentric) of an edge: the two tests may
            # legitimately disagree there
            e1 = tris[i, 1].astype(np.float64) - tris[i, 0]
            e2 = tris[i, 2].astype(np.float64) - tris[i, 0]
            P = org[i] + t64 * d.astype(np.float64) - tris[i, 0]
            M = np.stack([e1, e2], 1)
            uv = np.linalg.lstsq(M, P, rcond=None)[0]
            if min(uv[0], uv[1], 1 - uv[0] - uv[1]) < 1e-4:
                continue
            assert h["primID"][0] == 0, i
            assert abs(h["t"][0] - t64) <= 1e-5 * max(1.0, t64), (i, h["t"][0], t64)
            checked += 1
        elif t64 < 0:
            assert h["primID"][0] in (-1, 0)
    assert checked > 200
    del s


def test_sphere_closed_form():
    c = np.array([[1.0, 2.0, 3.0]], np.float32)
    r = np.array([0.75], np.float32)
    s = refcpu.Scene([dict(kind=1, vertices=c, radii=r)])
    rays = np.array([
        [1, 2, -5, 0, 0, 0, 1, 1e30],        # head-on: t = 8 - 0.75
        [1, 2, 3, 0, 0, 0, 1, 1e30],         # from the center: far root = r
        [1, 2.74, -5, 0, 0, 0, 1, 1e30],     # near-grazing
        [1, 2.76, -5, 0, 0, 0, 1, 1e30],     # miss
        [1, 2, -5, 0, 0, 0, 1, 7.0],         # tmax before hit
    ], np.float32)
    h = s.trace(rays)
    assert abs(h["t"][0] - 7.25) < 1e-5
    assert abs(h["t"][1] - 0.75) < 1e-6
    dy = 0.74
    assert abs(h["t"][2] - (8 - math.sqrt(0.75 ** 2 - dy ** 2))) < 2e-4
    assert h["primID"][3] == -1 and h["primID"][4] == -1
    # zero direction components are traced as 1e-6 (traceRays.dev.cu:42-44)
    np.testing.assert_allclose(h["N"][0], [0, 0, -1], atol=3e-5)


def _cyl_t_f64(o, d, a, b, r):
    """Capped cylinder, double precision: min t over tube and cap hits."""
    o, d, a, b = (np.asarray(x, np.float64) for x in (o, d, a, b))
    ax = b - a
    L_ = np.linalg.norm(ax)
    w = ax / L_
    best = math.inf
    oc = o - a
    dp = d - np.dot(d, w) * w
    op = oc - np.dot(oc, w) * w
    A_, B_, C_ = np.dot(dp, dp), 2 * np.dot(dp, op), np.dot(op, op) - r * r
    disc = B_ * B_ - 4 * A_ * C_
    if A_ > 0 and disc >= 0:
        for t in ((-B_ - math.sqrt(disc)) / (2 * A_), (-B_ + math.sqrt(disc)) / (2 * A_)):
            h = np.dot(oc + t * d, w)
            if t > 0 and 0 <= h <= L_:
                best = min(best, t)
    for base in (a, b):
        dn = np.dot(d, w)
        if abs(dn) > 1e-12:
            t = np.dot(base - o, w) / dn
            p = o + t * d - base
            if t > 0 and np.dot(p, p) <= r * r:
                best = min(best, t)
    return best


def test_cylinder_vs_f64():
    rng = np.random.default_rng(5)
    ok = 0
    for k in range(400):
        a = rng.uniform(-1, 1, 3).astype(np.float32)
        b = (a + rng.uniform(-2, 2, 3)).astype(np.float32)
        r = np.float32(rng.uniform(0.1, 0.6))
        s = refcpu.Scene([dict(kind=2, vertices=np.stack([a, b]), indices=np.array([[0, 1]], np.int32),
                               radii=np.array([r], np.float32))])
        o = rng.uniform(-4, 4, 3).astype(np.float32)
        tgt = (0.5 * (a + b) + rng.uniform(-0.5, 0.5, 3)).astype(np.float32)
        d = (tgt - o) / np.linalg.norm(tgt - o)
        d = d.astype(np.float32)
        h = s.trace(np.array([[*o, 0, *d, 1e30]], np.float32))
        t64 = _cyl_t_f64(o, d, a, b, float(r))
        if math.isinf(t64):
            continue
        if h["primID"][0] == 0:
            assert abs(h["t"][0] - t64) <= 1e-4 * max(1.0, t64), (k, h["t"][0], t64)
            ok += 1
    assert ok > 150


def test_bvh_equals_brute_force_and_ties():
    g = small_sphere_scene(24) + user_prim_scene(80, 80)
    s = refcpu.Scene(g)
    rays = random_rays(8000, (0, 0, 0), (30, 30, 30), seed=3)
    a = s.trace(rays)
    b = s.trace(rays, brute=True)
    for k in ("primID", "geomID"):
        assert np.array_equal(a[k], b[k])
    assert np.array_equal(a["t"].view(np.uint32), b["t"].view(np.uint32))
    assert a["primTests"] < b["primTests"] / 20
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    t = np.array([[0, 1, 2]], np.int32)
    dup = refcpu.Scene([dict(kind=0, vertices=v, indices=t), dict(kind=0, vertices=v, indices=t)])
    h = dup.trace(np.array([[0.2, 0.2, -1, 0, 0, 0, 1, 1e30]], np.float32))
    assert h["geomID"][0] == 0 and h["t"][0] == 1.0


# ---------------------------------------------------------------------------
# Camera: fovy / aspect closed forms
# ---------------------------------------------------------------------------

def test_camera_geometry():
    import ctypes as C
    W, H = 64, 48
    cam = refcpu.camera((1, 2, 3), (0, 0, 5), up=(0, 1, 0), fovy=60)
    p = refcpu.render_params(W, H)
    org, d = np.zeros(3, np.float32), np.zeros(3, np.float32)
    seed = C.c_uint64(0)
    # pixel (W/2 - 0.5, H/2 - 0.5) does not exist; use the bottom edge row 0
    L.ref_camera_ray(C.byref(cam), C.byref(p), W // 2, 0, org.ctypes.data, d.ctypes.data, C.byref(seed))
    np.testing.assert_allclose(org, [1, 2, 3])
    v = (0 + 0.5) / H - 0.5
    u = (W // 2 + 0.5) / W - 0.5
    f = 1 / (2 * math.tan(math.radians(30)))
    # dir_du = normalize(dir x up) = (-1,0,0) for dir=+z, up=+y; dir_dv = +y
    exp = np.array([-u * W / H, v, f])
    np.testing.assert_allclose(d, exp / np.linalg.norm(exp), atol=1e-6)


# ---------------------------------------------------------------------------
# Volume: trilinear, transfer function, majorants, DDA, Woodcock
# ---------------------------------------------------------------------------

def _vol(field, cpm=4, base=1.0, tf=None):
    v = refcpu.Volume(field, origin=(0.5, -1.0, 2.0), spacing=(0.5, 0.25, 1.0), cells_per_mc=cpm)
    if tf is None:
        x = np.linspace(0, 1, 16)
        tf = np.stack([x, 1 - x, 0.5 + 0 * x, x ** 2], -1)
    v.set_tf(tf, (0.0, 1.0), base)
    return v


def test_trilinear_reproduces_linear_fields_and_nan_outside():
    nz, ny, nx = 6, 7, 9
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    field = (0.1 * x + 0.03 * y + 0.07 * z).astype(np.float32)
    v = _vol(field)
    rng = np.random.default_rng(1)
    for _ in range(200):
        g = rng.uniform([0, 0, 0], [nx - 1.001, ny - 1.001, nz - 1.001])
        P = np.array([0.5, -1.0, 2.0]) + g * np.array([0.5, 0.25, 1.0])
        exp = 0.1 * g[0] + 0.03 * g[1] + 0.07 * g[2]
        assert abs(v.sample(P.astype(np.float32)) - exp) < 2e-5
    assert math.isnan(v.sample(np.array([0.4, 0.0, 3.0], np.float32)))
    assert math.isnan(v.sample(np.array([0.5 + 0.5 * (nx - 1), 0.0, 3.0], np.float32)))  # rel == numCells


def test_tf_map_and_majorant():
    x = np.linspace(0, 1, 16)
    tf = np.stack([x, 1 - x, 0.5 + 0 * x, np.abs(np.sin(7 * x))], -1).astype(np.float32)
    v = _vol(np.zeros((3, 3, 3), np.float32), base=2.5, tf=tf)
    rng = np.random.default_rng(2)
    for s in rng.uniform(-0.2, 1.2, 300):
        sc = min(max(s, 0.0), 1.0)
        exp = [np.interp(sc, x, tf[:, k]) for k in range(4)]
        exp[3] *= 2.5
        np.testing.assert_allclose(v.tf_map(s), exp, rtol=1e-5, atol=1e-6)
    for _ in range(200):
        lo, hi = sorted(rng.uniform(-0.1, 1.1, 2))
        m = v.tf_majorant(lo, hi)
        dense = max(v.tf_map(s)[3] for s in np.linspace(lo, hi, 400))
        assert m >= dense - 1e-5
        # the maximum of a piecewise-linear function is at an end or a knot
        knots = [v.tf_map(s)[3] for s in [lo, hi] + [k for k in x if lo < k < hi]]
        assert abs(m - max(knots)) < 1e-5
    assert v.tf_majorant(0.6, 0.4) == 0.0


def _brute_cells(org, d, tmax, grid):
    """Cells whose closed boxes the segment [max(0,t_enter), min(tmax,t_exit)]
    passes through (dense sampling + box tests, float64)."""
    org, d = np.asarray(org, np.float64), np.asarray(d, np.float64)
    t0, t1 = 0.0, tmax
    for k in range(3):
        if d[k] != 0:
            a, b = (0 - org[k]) / d[k], (grid[k] - org[k]) / d[k]
            t0, t1 = max(t0, min(a, b)), min(t1, max(a, b))
    if t0 > t1:
        return set()
    ts = np.linspace(t0, t1, 20001)[1:-1]
    P = org + ts[:, None] * d
    c = np.floor(P).astype(int)
    c = np.clip(c, 0, np.array(grid) - 1)
    return set(map(tuple, c))


def test_dda_against_brute_force():
    rng = np.random.default_rng(4)
    grid = (5, 4, 6)
    for k in range(300):
        org = rng.uniform(-3, 9, 3).astype(np.float32)
        tgt = rng.uniform(0, 1, 3) * np.array(grid)
        d = (tgt - org).astype(np.float32)
        cells, ts, n = refcpu.dda_cells(org, d, 1e30, grid)
        got = set(map(tuple, cells.tolist()))
        exp = _brute_cells(org, d, 1e30, grid)
        # the DDA may add cells touched only at a corner/edge (measure zero)
        assert exp <= got, (k, exp - got)
        assert len(got - exp) <= 2
        assert np.all(np.diff(ts[:, 0]) >= 0)          # monotone cell entries
        assert np.all(ts[1:, 0] == ts[:-1, 1]) or n < 2  # contiguous segments


def test_woodcock_homogeneous_transmittance():
    # constant density sigma: P(no collision within distance d) = exp(-sigma d)
    sigma, n = 0.35, 12
    field = np.full((n + 1, n + 1, n + 1), 0.7, np.float32)
    tf = np.array([[1, 1, 1, 0.5], [1, 1, 1, 0.5]], np.float32)
    v = refcpu.Volume(field, origin=(0, 0, 0), spacing=(1, 1, 1), cells_per_mc=4)
    v.set_tf(tf, (0.0, 1.0), sigma / 0.5)
    m = 40000
    rays = np.zeros((m, 8), np.float32)
    rays[:, 0:3] = (0.5, 3.3, -2.0)
    rays[:, 6] = 1.0
    rays[:, 7] = 1e30
    seeds = np.arange(m, dtype=np.uint64) * np.uint64(2654435761) + np.uint64(1)
    out, _ = v.sample_rays(rays, seeds, exact_log=True)
    passed = np.isinf(out[:, 0]).mean()
    exp = math.exp(-sigma * n)
    assert abs(passed - exp) < 4 * math.sqrt(exp * (1 - exp) / m) + 1e-4
    t = out[np.isfinite(out[:, 0]), 0] - 2.0  # distance into the medium
    # collision distances are exponential(sigma) truncated at n
    assert abs(np.mean(t) - (1 / sigma - n * math.exp(-sigma * n) / (1 - math.exp(-sigma * n)))) < 0.05


# ---------------------------------------------------------------------------
# Shading chain: Lambertian plane under uniform ambient light, closed form
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("rr", [0, 1])
def test_lambert_plane_furnace_closed_form(rr):
    # big plane y=0 facing +y, camera looking down; one diffuse bounce then
    # escape. E[pixel] = E_unif[4 a alb mu^2/(1+4mu)] + E_cos[pi alb a 4mu/(4mu+1)]
    # with alb = 0.85 c / pi (AnariMatte) -- see shadeRays.cu:676-693, 833-871.
    # rr = 1: Russian roulette on the bounced ray (survival q = max(tp) =
    # 0.85 c, weight 1/q) leaves the expectation unchanged; a survivor
    # without the 1/q weight lowers the bounce term by the factor q.
    c, amb = 0.6, 0.8
    S = 1e4
    v = np.array([[-S, 0, -S], [S, 0, -S], [S, 0, S], [-S, 0, S]], np.float32)
    t = np.array([[0, 2, 1], [0, 3, 2]], np.int32)
    s = refcpu.Scene([dict(kind=0, vertices=v, indices=t, color=(c, c, c))])
    W, H, spp = 40, 30, 16 if rr == 0 else 64
    acc, _, _ = s.render(refcpu.camera((0, 10, 0), (0.01, -1, 0.02), fovy=30),
                         refcpu.render_params(W, H, spp=spp, ambient=amb, russian_roulette=rr))
    pix = acc[..., 0].ravel() / spp
    alb = 0.85 * c / math.pi
    mu = np.linspace(0, 1, 200001)
    i1 = 4 * amb * alb * np.trapezoid(mu ** 2 / (1 + 4 * mu), mu)
    i2 = math.pi * alb * amb * np.trapezoid(4 * mu / (4 * mu + 1) * 2 * mu, mu)
    exp = i1 + i2
    se = pix.std() / math.sqrt(len(pix))
    assert abs(pix.mean() - exp) < 4 * se + 1e-3, (pix.mean(), exp, se)
    assert np.allclose(acc[..., 3], spp)  # alpha: every primary ray hits


def _open_box():
    """Five matte quads of a 10-unit box, open at the top (y = 10)."""
    q = [[(0, 0, 0), (10, 0, 0), (10, 0, 10), (0, 0, 10)],      # floor
         [(0, 0, 0), (0, 10, 0), (10, 10, 0), (10, 0, 0)],      # z = 0
         [(0, 0, 10), (10, 0, 10), (10, 10, 10), (0, 10, 10)],  # z = 10
         [(0, 0, 0), (0, 0, 10), (0, 10, 10), (0, 10, 0)],      # x = 0
         [(10, 0, 0), (10, 10, 0), (10, 10, 10), (10, 0, 10)]]  # x = 10
    v = np.array(q, np.float32).reshape(-1, 3)
    t = np.array([[4 * k, 4 * k + 1, 4 * k + 2] for k in range(5)] +
                 [[4 * k, 4 * k + 2, 4 * k + 3] for k in range(5)], np.int32)
    return refcpu.Scene([dict(kind=0, vertices=v, indices=t, color=(0.8, 0.75, 0.7))])


def test_russian_roulette_is_unbiased_over_many_bounces():
    """Paths bounce up to 8 times inside an open box (up to 8 diffuse
    bounces; the only light is the ambient sky seen through the opening).
    Russian roulette from depth 2 kills paths with probability 1 - q, q =
    min(max(tp), 0.95), and weights survivors by 1/q: the expected image is
    the one without roulette. Per band of rows, the mean over 32 independent
    frames with roulette equals the mean without within 3.5 standard errors;
    roulette does end paths early (fewer rays). Dropping the 1/q weight, or
    killing with probability q instead of 1 - q, moves the bands by many
    standard errors."""
    s = _open_box()
    cam = refcpu.camera((5.0, 24.0, -6.0), (0.0, -18.0, 11.0), fovy=35)
    W, H, spp, frames = 24, 18, 8, 32

    def frame_bands(rr, first):
        out, rays = [], 0
        for f in range(frames):
            acc, _, n = s.render(cam, refcpu.render_params(W, H, spp=spp, accum_id=first + f * spp, ambient=1.0,
                                                            max_depth=8, max_diffuse=8, russian_roulette=rr))
            rays += n
            px = acc[..., :3].sum(-1) / spp
            out.append(px.reshape(6, -1).mean(1))  # six bands of three rows
        return np.array(out, np.float64), rays

    on, rays_on = frame_bands(2, 0)
    off, rays_off = frame_bands(0, 100_000)
    z = (on.mean(0) - off.mean(0)) / np.sqrt(on.var(0, ddof=1) / frames + off.var(0, ddof=1) / frames)
    print(f"bands off {np.round(off.mean(0), 4)} z {np.round(z, 2)} rays {rays_on} vs {rays_off}")
    assert np.all(np.abs(z) <= 3.5), z
    assert off.mean() > 0.1
    assert rays_on < 0.9 * rays_off


@pytest.mark.parametrize("cpm", [4, 3])
def test_structured_macrocells_bound_every_sample(cpm):
    """StructuredData_computeMCs (StructuredData.cu:30-61) + mapMCs: the
    range of macrocell c spans the (cpm+1)^3 voxels its cells touch, so every
    trilinear sample taken inside c lies in c's range, and c's majorant is
    at least the transfer function's opacity at every such sample. A white-
    noise field makes the face voxels matter (a cpm^3 corner range misses
    them) and a transfer function peaked between its knots makes interior
    knots matter (a majorant of the range's end values misses the peak).
    cpm = 3 does not divide the 20 cells: the last macrocell is partial."""
    n = 21
    field = np.random.default_rng(5).random((n, n, n)).astype(np.float32)
    origin, spacing = np.array([0.5, -1.0, 2.0]), np.array([0.5, 0.25, 1.0])
    v = refcpu.Volume(field, origin=tuple(origin), spacing=tuple(spacing), cells_per_mc=cpm)
    x = np.linspace(0, 1, 9)
    tf = np.stack([x, 1 - x, 0.5 + 0 * x, np.where(np.arange(9) == 4, 1.0, 0.1)], -1).astype(np.float32)
    v.set_tf(tf, (0.0, 1.0), 2.0)
    ranges, maj = v.macrocells()
    mx, my, mz = v.mc_dims()
    assert (mx, my, mz) == ((n - 1 + cpm - 1) // cpm,) * 3
    rng = np.random.default_rng(6)
    hi = origin + spacing * (n - 1)
    P = rng.uniform(origin, hi, (6000, 3))
    # and points on and next to macrocell faces (x = k cpm cells)
    k = rng.integers(1, mx, 3000)
    Q = rng.uniform(origin, hi, (3000, 3))
    Q[:, 0] = origin[0] + spacing[0] * cpm * k + rng.choice([-1e-4, 0.0, 1e-4], 3000)
    P = np.concatenate([P, Q]).astype(np.float32)
    bad_range = bad_maj = 0
    for p in P:
        s = v.sample(p)
        if np.isnan(s):
            continue
        c = np.minimum(((p.astype(np.float64) - origin) / (spacing * cpm)).astype(int), [mx - 1, my - 1, mz - 1])
        lo_, hi_ = ranges[c[2], c[1], c[0]]
        bad_range += not (lo_ <= s <= hi_)
        bad_maj += not (v.tf_map(s)[3] <= maj[c[2], c[1], c[0]])
    assert bad_range == 0 and bad_maj == 0, (bad_range, bad_maj)
    # the majorant is the transfer function's maximum over the range, times
    # the base density (TransferFunction.h:92-116): the peak knot when the
    # range covers 0.5
    cover = (ranges[..., 0] <= 0.5) & (ranges[..., 1] >= 0.5)
    assert cover.any() and np.allclose(maj[cover], 2.0)
