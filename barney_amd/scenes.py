"""Deterministic synthetic scenes for the five benchmark workloads.

Random numbers come from SplitMix64 in counter mode (value i of stream s is
splitmix64(seed ^ s * phi + (i + 1) * phi)), vectorized with numpy, so every
array is reproducible from (seed, stream). Seeds are 0x5EED0000 + workload
index.
"""
import numpy as np

PHI = np.uint64(0x9E3779B97F4A7C15)


def _splitmix(x):
    x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return x ^ (x >> np.uint64(31))


def uniform(seed, stream, n, lo=0.0, hi=1.0):
    """n uniform floats in [lo, hi) from stream 'stream' of 'seed' (float64)."""
    with np.errstate(over="ignore"):
        base = np.uint64(seed) ^ (np.uint64(stream) * PHI)
        i = np.arange(1, n + 1, dtype=np.uint64)
        v = _splitmix(base + i * PHI)
    u = (v >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return lo + (hi - lo) * u


# ---------------------------------------------------------------------------
# Workload 1: 1,024 spheres tessellated to 512 triangles each
# ---------------------------------------------------------------------------

def sphere_mesh(center, radius, stacks=16, slices=16):
    """Lat-long sphere: (stacks+1)*(slices+1) vertices, 2*stacks*slices
    triangles with outward winding (pole triangles are degenerate, kept)."""
    th = np.linspace(0.0, np.pi, stacks + 1)
    ph = np.linspace(0.0, 2.0 * np.pi, slices + 1)
    T, P = np.meshgrid(th, ph, indexing="ij")
    x = np.sin(T) * np.cos(P)
    y = np.cos(T)
    z = np.sin(T) * np.sin(P)
    verts = np.stack([x, y, z], -1).reshape(-1, 3) * radius + np.asarray(center)
    i, j = np.meshgrid(np.arange(stacks), np.arange(slices), indexing="ij")
    v00 = i * (slices + 1) + j
    v01 = v00 + 1
    v10 = v00 + (slices + 1)
    v11 = v10 + 1
    t0 = np.stack([v00, v01, v11], -1).reshape(-1, 3)
    t1 = np.stack([v00, v11, v10], -1).reshape(-1, 3)
    tris = np.stack([t0, t1], 1).reshape(-1, 3)
    return verts.astype(np.float32), tris.astype(np.int32)


def workload1(num_spheres=1024, stacks=16, slices=16, seed=0x5EED0000):
    """Returns dict(geoms=[(verts, tris, color)], camera=..., size=(1024,768))."""
    c = uniform(seed, 1, 3 * num_spheres, 0.0, 100.0).reshape(-1, 3)
    r = uniform(seed, 2, num_spheres, 0.5, 2.0)
    col = uniform(seed, 3, 3 * num_spheres, 0.2, 0.9).reshape(-1, 3)
    geoms = []
    for s in range(num_spheres):
        v, t = sphere_mesh(c[s], r[s], stacks, slices)
        geoms.append((v, t, col[s].astype(np.float32)))
    cam = dict(position=(-50.0, 50.0, -50.0), direction=(100.0, 0.0, 100.0), up=(0.0, 1.0, 0.0), fovy=60.0)
    return dict(geoms=geoms, camera=cam, size=(1024, 768))


# ---------------------------------------------------------------------------
# Workload 2: structured volume, Gaussian blobs + value noise
# ---------------------------------------------------------------------------

def _value_noise(n, lattice, seed, stream):
    """Trilinear value noise over a (lattice+1)^3 random lattice, evaluated on
    an n^3 voxel grid via separable 1-D interpolation matrices."""
    lat = uniform(seed, stream, (lattice + 1) ** 3).reshape(lattice + 1, lattice + 1, lattice + 1)
    x = np.arange(n, dtype=np.float64) * (lattice / max(n - 1, 1))
    i0 = np.minimum(np.floor(x).astype(np.int64), lattice - 1)
    f = x - i0
    M = np.zeros((n, lattice + 1))
    M[np.arange(n), i0] = 1.0 - f
    M[np.arange(n), i0 + 1] += f
    # out[z,y,x] = sum M[z,k] M[y,j] M[x,i] lat[k,j,i]
    tmp = np.tensordot(M, lat, axes=([1], [0]))            # z, j, i
    tmp = np.tensordot(tmp, M, axes=([2], [1]))            # z, j, x
    tmp = np.einsum("yj,zjx->zyx", M, tmp, optimize=True)  # z, y, x
    return tmp


def structured_field(n=512, blobs=32, sigma=(10.0, 60.0), octaves=4, seed=0x5EED0001, dtype=np.float32):
    """n^3 field (z, y, x) = sum of Gaussian blobs + value noise, in [0, 1]."""
    centers = uniform(seed, 1, 3 * blobs, 0.0, float(n)).reshape(-1, 3)
    sig = uniform(seed, 2, blobs, sigma[0] * n / 512.0, sigma[1] * n / 512.0)
    amp = uniform(seed, 3, blobs, 0.5, 1.0)
    ax = np.arange(n, dtype=np.float32)
    field = np.zeros((n, n, n), dtype=np.float32)
    for b in range(blobs):
        s2 = np.float32(2.0 * sig[b] * sig[b])
        gx = np.exp(-((ax - np.float32(centers[b, 0])) ** 2) / s2)
        gy = np.exp(-((ax - np.float32(centers[b, 1])) ** 2) / s2)
        gz = np.exp(-((ax - np.float32(centers[b, 2])) ** 2) / s2) * np.float32(amp[b])
        field += gz[:, None, None] * (gy[:, None] * gx[None, :])[None, :, :]
    amp_n = 0.5
    for o in range(octaves):
        lattice = 8 << o
        field += (amp_n * _value_noise(n, lattice, seed, 10 + o)).astype(np.float32)
        amp_n *= 0.5
    lo, hi = float(field.min()), float(field.max())
    field = (field - lo) / max(hi - lo, 1e-30)
    return field.astype(dtype)


def ramp_tf(n=256, alpha_scale=1.0):
    """256-entry RGBA piecewise-linear ramp: blue -> white -> red, alpha
    rising linearly from 0 with a dead zone over the lowest 20 percent."""
    x = np.linspace(0.0, 1.0, n)
    r = np.interp(x, [0.0, 0.5, 1.0], [0.1, 0.9, 1.0])
    g = np.interp(x, [0.0, 0.5, 1.0], [0.2, 0.9, 0.2])
    b = np.interp(x, [0.0, 0.5, 1.0], [1.0, 0.9, 0.1])
    a = np.clip((x - 0.2) / 0.8, 0.0, 1.0) * alpha_scale
    return np.stack([r, g, b, a], -1).astype(np.float32)


def workload2(n=512, seed=0x5EED0001):
    field = structured_field(n=n, seed=seed)
    c = (n - 1) / 2.0
    pos = (c - 0.9 * n, c + 0.5 * n, c - 1.6 * n)
    cam = dict(position=pos, direction=(c - pos[0], c - pos[1], c - pos[2]), up=(0.0, 1.0, 0.0), fovy=40.0)
    return dict(field=field, tf=ramp_tf(), base_density=0.05, cells_per_mc=16, camera=cam, size=(1920, 1080),
                spp=16)


# ---------------------------------------------------------------------------
# Workload 4: differentiable emission-absorption volume
# ---------------------------------------------------------------------------

def random_tf(seed, n=256, knots=9):
    """Smooth random RGBA transfer function (piecewise-linear through
    'knots' random control points), opacity in [0, 0.6]."""
    x = np.linspace(0.0, 1.0, n)
    kx = np.linspace(0.0, 1.0, knots)
    vals = uniform(seed, 7, 4 * knots).reshape(knots, 4)
    vals[:, 3] *= 0.6
    vals[0, 3] = 0.0
    return np.stack([np.interp(x, kx, vals[:, k]) for k in range(4)], -1).astype(np.float32)


def orbit_cameras(center, radius, count, fovy):
    """'count' views on a Fibonacci sphere around 'center', looking at it."""
    cams = []
    ga = np.pi * (3.0 - np.sqrt(5.0))
    for k in range(count):
        y = 1.0 - 2.0 * (k + 0.5) / count
        r = np.sqrt(max(0.0, 1.0 - y * y))
        a = ga * k
        d = np.array([np.cos(a) * r, y, np.sin(a) * r])
        pos = np.asarray(center) + radius * d
        up = (0.0, 1.0, 0.0) if abs(y) < 0.99 else (1.0, 0.0, 0.0)
        cams.append(dict(position=tuple(pos), direction=tuple(np.asarray(center) - pos), up=up, fovy=fovy))
    return cams


def workload4(n=256, views=64, size=512, seed=0x5EED0003):
    dens = structured_field(n=n, blobs=32, sigma=(10.0 * n / 512 * 2, 60.0 * n / 512 * 2), octaves=0, seed=seed)
    c = (n - 1) / 2.0
    diag = (n - 1) * np.sqrt(3.0)
    cams = orbit_cameras((c, c, c), 1.5 * diag, views, fovy=40.0)
    return dict(density=dens, tf_target=random_tf(seed + 1), tf_init=random_tf(seed + 2), cameras=cams,
                size=(size, size), step=0.5, base_density=0.05)


# ---------------------------------------------------------------------------
# Workload 3: batched multi-view mixed scene
# ---------------------------------------------------------------------------

def value_noise_2d(n, lattice, seed, stream):
    """Bilinear value noise on an n x n grid over a (lattice+1)^2 lattice."""
    lat = uniform(seed, stream, (lattice + 1) ** 2).reshape(lattice + 1, lattice + 1)
    x = np.arange(n, dtype=np.float64) * (lattice / max(n - 1, 1))
    i0 = np.minimum(np.floor(x).astype(np.int64), lattice - 1)
    f = x - i0
    M = np.zeros((n, lattice + 1))
    M[np.arange(n), i0] = 1.0 - f
    M[np.arange(n), i0 + 1] += f
    return M @ lat @ M.T


def heightfield(n=2236, height=120.0, seed=0x5EED0002):
    """n x n jittered vertex grid (x,z in [0, n-1], +-0.25 cell jitter),
    heights from 4-octave value noise; 2 (n-1)^2 triangles facing +y."""
    h = np.zeros((n, n))
    amp = 1.0
    for o in range(4):
        h += amp * value_noise_2d(n, 4 << o, seed, 20 + o)
        amp *= 0.5
    h = (h - h.min()) / max(h.max() - h.min(), 1e-12) * height
    jx = uniform(seed, 30, n * n, -0.25, 0.25).reshape(n, n)
    jz = uniform(seed, 31, n * n, -0.25, 0.25).reshape(n, n)
    zz, xx = np.meshgrid(np.arange(n, dtype=np.float64), np.arange(n, dtype=np.float64), indexing="ij")
    verts = np.stack([xx + jx, h, zz + jz], -1).reshape(-1, 3).astype(np.float32)
    i, j = np.meshgrid(np.arange(n - 1), np.arange(n - 1), indexing="ij")
    v00 = (i * n + j).ravel()
    v01 = v00 + 1
    v10 = v00 + n
    v11 = v10 + 1
    # winding so that cross(v1-v0, v2-v0) points up (+y)
    t0 = np.stack([v00, v10, v11], -1)
    t1 = np.stack([v00, v11, v01], -1)
    tris = np.stack([t0, t1], 1).reshape(-1, 3).astype(np.int32)
    return verts, tris, h


def split_material(n, seed, stream, fractions=(0.6, 0.3, 0.1)):
    """Random assignment of n items to len(fractions) classes."""
    u = uniform(seed, stream, n)
    edges = np.cumsum(fractions)
    return np.searchsorted(edges, u, side="right").clip(0, len(fractions) - 1)


def workload3(grid=2236, spheres=200_000, cylinders=100_000, views=64, seed=0x5EED0002):
    """Mixed scene: heightfield triangles + spheres + cylinders, materials
    diffuse/glossy/dielectric 60/30/10, 64 views of uneven size."""
    verts, tris, h = heightfield(grid, seed=seed)
    ext = float(grid - 1)
    hmax = float(h.max())
    # spheres and cylinders scattered over the terrain, above the surface
    sx = uniform(seed, 40, spheres, 0.0, ext)
    sz = uniform(seed, 41, spheres, 0.0, ext)
    gi = np.clip(sz.astype(np.int64), 0, grid - 1) * grid + np.clip(sx.astype(np.int64), 0, grid - 1)
    sr = uniform(seed, 42, spheres, 0.5, 3.0)
    sy = h.ravel()[gi] + sr + uniform(seed, 43, spheres, 0.0, 20.0)
    centers = np.stack([sx, sy, sz], -1).astype(np.float32)
    cx = uniform(seed, 50, cylinders, 0.0, ext)
    cz = uniform(seed, 51, cylinders, 0.0, ext)
    ci = np.clip(cz.astype(np.int64), 0, grid - 1) * grid + np.clip(cx.astype(np.int64), 0, grid - 1)
    cy = h.ravel()[ci] + uniform(seed, 52, cylinders, 0.0, 10.0)
    a = np.stack([cx, cy, cz], -1)
    d = uniform(seed, 53, 3 * cylinders, -1.0, 1.0).reshape(-1, 3)
    d /= np.linalg.norm(d, axis=1, keepdims=True) + 1e-9
    length = uniform(seed, 54, cylinders, 2.0, 10.0)
    b = a + d * length[:, None]
    cverts = np.concatenate([a, b]).astype(np.float32)
    cidx = np.stack([np.arange(cylinders), np.arange(cylinders) + cylinders], 1).astype(np.int32)
    crad = uniform(seed, 55, cylinders, 0.1, 0.5).astype(np.float32)
    # materials: 0 diffuse, 1 glossy, 2 dielectric
    mats = []
    for k in range(4):
        col = uniform(seed, 60 + k, 3, 0.2, 0.9)
        mats.append(dict(type="matte", color=tuple(col)))
    for k in range(4):
        col = uniform(seed, 70 + k, 3, 0.2, 0.9)
        rough = float(uniform(seed, 80 + k, 1, 0.05, 0.5)[0])
        mats.append(dict(type="pbr", color=tuple(col), roughness=rough, metallic=0.0))
    mats.append(dict(type="glass", ior=1.5))
    def pick(cls, n, stream):
        sub = uniform(seed, stream, n)
        out = np.empty(n, np.int64)
        out[cls == 0] = (sub[cls == 0] * 4).astype(np.int64)
        out[cls == 1] = 4 + (sub[cls == 1] * 4).astype(np.int64)
        out[cls == 2] = 8
        return out
    geoms = []
    tcls = split_material(len(tris), seed, 90)
    tmat = pick(tcls, len(tris), 91)
    for m in range(len(mats)):
        sel = tmat == m
        if sel.any():
            geoms.append(dict(kind=0, vertices=verts, indices=tris[sel], material=m))
    scls = split_material(spheres, seed, 92)
    smat = pick(scls, spheres, 93)
    for m in range(len(mats)):
        sel = smat == m
        if sel.any():
            geoms.append(dict(kind=1, vertices=centers[sel], radii=sr[sel].astype(np.float32), material=m))
    ccls = split_material(cylinders, seed, 94)
    cmat = pick(ccls, cylinders, 95)
    for m in range(len(mats)):
        sel = cmat == m
        if sel.any():
            idx = np.nonzero(sel)[0]
            geoms.append(dict(kind=2, vertices=cverts, indices=cidx[idx], radii=crad[idx], material=m))
    # 64 views of uneven resolution looking down at the terrain
    cams, widths, heights = [], [], []
    wv = (256 + np.floor(uniform(seed, 100, views) * 769)).astype(int)
    hv = (256 + np.floor(uniform(seed, 101, views) * 769)).astype(int)
    px = uniform(seed, 102, views, 0.1 * ext, 0.9 * ext)
    pz = uniform(seed, 103, views, 0.1 * ext, 0.9 * ext)
    # camera heights and look-at offsets scale with the terrain (1 at the
    # full 2236^2 grid, so reduced test scenes still look at geometry)
    sc = ext / 2235.0
    py = hmax + uniform(seed, 104, views, 40.0, 300.0) * sc
    tx = px + uniform(seed, 105, views, -300.0, 300.0) * sc
    tz = pz + uniform(seed, 106, views, -300.0, 300.0) * sc
    for k in range(views):
        pos = (float(px[k]), float(py[k]), float(pz[k]))
        tgt = (float(tx[k]), float(0.5 * hmax), float(tz[k]))
        cams.append(dict(position=pos, direction=tuple(tgt[i] - pos[i] for i in range(3)), fovy=50.0))
        widths.append(int(wv[k]))
        heights.append(int(hv[k]))
    return dict(geoms=geoms, materials=mats, cameras=cams, widths=widths, heights=heights, spp=4, max_depth=8,
                rr_depth=2, max_diffuse=8)


def product_materials(api, mats):
    out = []
    for m in mats:
        if m["type"] == "matte":
            out.append(api.matte(m["color"], opacity=m.get("opacity", 1.0)))
        elif m["type"] == "pbr":
            out.append(api.pbr(m["color"], metallic=m.get("metallic", 0.0), roughness=m["roughness"],
                               transmission=m.get("transmission", 0.0), opacity=m.get("opacity", 1.0)))
        else:
            out.append(api.pbr((1.0, 1.0, 1.0), transmission=1.0, ior=m["ior"]))
    return out


# ---------------------------------------------------------------------------
# Workload 5: unstructured tetrahedral mesh
# ---------------------------------------------------------------------------

# corner c of a hex = (c & 1, (c >> 1) & 1, (c >> 2) & 1); two mirror-image
# 5-tet splits, alternated by cell parity so shared faces use the same
# diagonal (a conforming mesh)
_TETS_EVEN = ((0, 1, 2, 4), (3, 1, 7, 2), (5, 1, 4, 7), (6, 2, 7, 4), (1, 2, 4, 7))
_TETS_ODD = ((1, 0, 5, 3), (2, 0, 3, 6), (4, 0, 6, 5), (7, 3, 5, 6), (0, 3, 6, 5))


def tet_grid(n=128, jitter=0.2, seed=0x5EED0004, blobs=16):
    """(n+1)^3 vertex grid, jittered by +-jitter cells, each hex split into 5
    positively oriented tets (det(v1-v0, v2-v0, v3-v0) > 0). Per-vertex
    scalars: sum of Gaussian blobs normalized to [0, 1].
    Returns vertices (float32 [V,3]), scalars (float32 [V]), tets (int32 [T,4])."""
    m = n + 1
    zz, yy, xx = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
    P = np.stack([xx, yy, zz], -1).reshape(-1, 3).astype(np.float64)
    P += uniform(seed, 1, 3 * m ** 3, -jitter, jitter).reshape(-1, 3)
    centers = uniform(seed, 2, 3 * blobs, 0.0, float(n)).reshape(-1, 3)
    sig = uniform(seed, 3, blobs, 0.08 * n, 0.25 * n)
    amp = uniform(seed, 4, blobs, 0.5, 1.0)
    s = np.zeros(len(P))
    for b in range(blobs):
        d2 = ((P - centers[b]) ** 2).sum(1)
        s += amp[b] * np.exp(-d2 / (2.0 * sig[b] ** 2))
    s = (s - s.min()) / max(s.max() - s.min(), 1e-30)
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    base = (k * m + j) * m + i
    corner = [(c & 1) + ((c >> 1) & 1) * m + ((c >> 2) & 1) * m * m for c in range(8)]
    parity = ((i + j + k) & 1).ravel()
    base = base.ravel()
    tets = np.empty((n ** 3, 5, 4), np.int64)
    for t in range(5):
        for q in range(4):
            even = base + corner[_TETS_EVEN[t][q]]
            odd = base + corner[_TETS_ODD[t][q]]
            tets[:, t, q] = np.where(parity == 0, even, odd)
    tets = tets.reshape(-1, 4)
    # orient: swap the last two vertices where the triple product is negative
    a, b, c, d = (P[tets[:, q]] for q in range(4))
    det = np.einsum("ij,ij->i", np.cross(b - a, c - a), d - a)
    neg = det < 0
    tets[neg, 2], tets[neg, 3] = tets[neg, 3].copy(), tets[neg, 2].copy()
    return P.astype(np.float32), s.astype(np.float32), tets.astype(np.int32)


def workload5(n=128, seed=0x5EED0004):
    verts, scal, tets = tet_grid(n, seed=seed)
    c = n / 2.0
    pos = (c - 0.9 * n, c + 0.6 * n, c - 1.5 * n)
    cam = dict(position=pos, direction=(c - pos[0], c - pos[1], c - pos[2]), up=(0.0, 1.0, 0.0), fovy=45.0)
    return dict(vertices=verts, scalars=scal, tets=tets, tf=ramp_tf(), base_density=0.1, iso=0.5,
                iso_color=(0.9, 0.75, 0.4), camera=cam, size=(1024, 1024), spp=8)


# VTK hex corner order in the (x + 2y + 4z) corner numbering of tet_grid
_HEX_VTK = (0, 1, 3, 2, 4, 5, 7, 6)
# VTK hex faces (4 hex-local vertices each, cyclic)
_HEX_FACES = ((0, 1, 2, 3), (4, 5, 6, 7), (0, 1, 5, 4), (1, 2, 6, 5), (2, 3, 7, 6), (3, 0, 4, 7))
# 5-tet split of a VTK-ordered hex
_HEX_TETS5 = ((0, 1, 3, 4), (1, 2, 3, 6), (1, 4, 5, 6), (3, 4, 6, 7), (1, 3, 4, 6))

CELL_TET, CELL_HEX, CELL_PRISM, CELL_PYR = 10, 12, 13, 14
CELL_POLYHEDRON = 42


def face_stream(faces):
    """A polyhedron's index stream as UMeshField reads it: the number of
    faces, then each face's vertex count and vertex indices."""
    out = [len(faces)]
    for f in faces:
        out.append(len(f))
        out.extend(int(v) for v in f)
    return out


def mixed_cells(n=10, seed=0x5EED0005, blobs=6, polyhedra=False):
    """A conforming mixed mesh in UMeshField's cell layout: an n^3 lattice
    of cells under a shear (an affine map, so every face stays planar); cell
    (i, j, k) is, by (i + 2j + 3k) mod 4, one hex, two prisms (split through
    the bottom / top diagonal), six pyramids on a center vertex, or five
    positively oriented tets. Per-vertex scalars: Gaussian blobs in [0, 1].
    polyhedra=True writes every other hex (i + j + k even) as a polyhedron
    with its six quad faces and the first prism of each split as one with
    two triangles and three quads (VTK face streams).
    Returns vertices (float32 [V,3]), scalars (float32 [V]) and
    cells = (indices int32, offsets int32, types uint8)."""
    m = n + 1
    A = np.array([[1.0, 0.2, 0.1], [0.0, 1.0, 0.15], [0.05, 0.0, 1.0]])
    zz, yy, xx = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
    G = np.stack([xx, yy, zz], -1).reshape(-1, 3).astype(np.float64)
    ci, cj, ck = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    ci, cj, ck = ci.ravel(), cj.ravel(), ck.ravel()
    centers_g = np.stack([ci, cj, ck], -1) + 0.5
    Pg = np.concatenate([G, centers_g])  # lattice points, then cell centers
    P = Pg @ A.T
    bc = uniform(seed, 2, 3 * blobs, 0.0, float(n)).reshape(-1, 3) @ A.T
    sig = uniform(seed, 3, blobs, 0.15 * n, 0.4 * n)
    amp = uniform(seed, 4, blobs, 0.5, 1.0)
    s = np.zeros(len(P))
    for b in range(blobs):
        s += amp[b] * np.exp(-((P - bc[b]) ** 2).sum(1) / (2.0 * sig[b] ** 2))
    s = (s - s.min()) / max(s.max() - s.min(), 1e-30)
    corner = [(c & 1) + ((c >> 1) & 1) * m + ((c >> 2) & 1) * m * m for c in range(8)]
    indices, offsets, types = [], [], []

    def add(kind, vs):
        offsets.append(len(indices))
        types.append(kind)
        indices.extend(int(v) for v in vs)

    for c in range(n ** 3):
        i, j, k = ci[c], cj[c], ck[c]
        base = (k * m + j) * m + i
        h = [base + corner[_HEX_VTK[q]] for q in range(8)]
        kind = (i + 2 * j + 3 * k) % 4
        if kind == 0 and polyhedra and (i + j + k) % 2 == 0:
            add(CELL_POLYHEDRON, face_stream([[h[q] for q in f] for f in _HEX_FACES]))
        elif kind == 0:
            add(CELL_HEX, h)
        elif kind == 1 and polyhedra:
            add(CELL_POLYHEDRON, face_stream([(h[0], h[1], h[2]), (h[4], h[5], h[6]), (h[0], h[1], h[5], h[4]),
                                              (h[1], h[2], h[6], h[5]), (h[2], h[0], h[4], h[6])]))
            add(CELL_PRISM, (h[0], h[2], h[3], h[4], h[6], h[7]))
        elif kind == 1:
            add(CELL_PRISM, (h[0], h[1], h[2], h[4], h[5], h[6]))
            add(CELL_PRISM, (h[0], h[2], h[3], h[4], h[6], h[7]))
        elif kind == 2:
            apex = m ** 3 + c
            for f in _HEX_FACES:
                add(CELL_PYR, [h[q] for q in f] + [apex])
        else:
            for t in _HEX_TETS5:
                v = [h[q] for q in t]
                a, b, cc, d = (P[x] for x in v)
                if np.dot(np.cross(b - a, cc - a), d - a) < 0:
                    v[2], v[3] = v[3], v[2]
                add(CELL_TET, v)
    return (P.astype(np.float32), s.astype(np.float32),
            (np.array(indices, np.int32), np.array(offsets, np.int32), np.array(types, np.uint8)))


def deformed_elements(seed=0x5EED0006, jitter=0.2, spacing=3.0):
    """One pyramid, one prism and one hex, each a unit element with its
    vertices moved by up to +-jitter (bilinear faces, non-affine maps), set
    'spacing' apart along x so that no two overlap. Scalars: a smooth
    nonlinear function of position. Same return layout as mixed_cells."""
    unit = {CELL_PYR: [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0.5, 0.5, 1)],
            CELL_PRISM: [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (0, 1, 1)],
            CELL_HEX: [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]}
    P, indices, offsets, types = [], [], [], []
    for e, (kind, vs) in enumerate(unit.items()):
        off = len(P)
        jit = uniform(seed, 10 + e, 3 * len(vs), -jitter, jitter).reshape(-1, 3)
        for q, v in enumerate(vs):
            P.append(np.array(v, np.float64) + jit[q] + np.array([spacing * e, 0.0, 0.0]))
        offsets.append(len(indices))
        types.append(kind)
        indices.extend(range(off, off + len(vs)))
    P = np.array(P)
    s = 0.5 + 0.3 * np.sin(1.3 * P[:, 0] + 0.4) * np.cos(0.9 * P[:, 1]) + 0.2 * P[:, 2] ** 2
    return (P.astype(np.float32), s.astype(np.float32),
            (np.array(indices, np.int32), np.array(offsets, np.int32), np.array(types, np.uint8)))


def amr_field(n=32, block=8, seed=0x5EED0007, blobs=8):
    """A two-level block-structured AMR field (BlockStructuredField's
    layout): level 1 (cells 2 units wide, refinement 2) covers [0, 2n)^3
    with n^3 cells, level 0 (1 unit) covers the middle [n/2, 3n/2)^3 with
    n^3 cells, both in block^3-cell blocks; every cell holds a smooth
    Gaussian-blob field at its center, normalized to [0, 1].
    Returns origins, dims (int32 [B,3]), levels (int32 [B]), offsets
    (uint64 [B]), refinements (int32 [2]), scalars (float32)."""
    assert n % block == 0 and (n // 2) % block == 0
    nb = n // block
    bc = uniform(seed, 1, 3 * blobs, 0.2 * n, 1.8 * n).reshape(-1, 3)
    sig = uniform(seed, 2, blobs, 0.1 * n, 0.4 * n)
    amp = uniform(seed, 3, blobs, 0.5, 1.0)

    def field(P):
        s = np.zeros(len(P))
        for b in range(blobs):
            s += amp[b] * np.exp(-((P - bc[b]) ** 2).sum(1) / (2.0 * sig[b] ** 2))
        return s

    origins, levels, vals = [], [], []
    c = np.arange(block)
    cz, cy, cx = np.meshgrid(c, c, c, indexing="ij")
    local = np.stack([cx.ravel(), cy.ravel(), cz.ravel()], -1)  # x fastest
    for level, cs, base in ((1, 2.0, 0), (0, 1.0, n // 2)):
        for k in range(nb):
            for j in range(nb):
                for i in range(nb):
                    o = np.array([base + i * block, base + j * block, base + k * block])
                    origins.append(o)
                    levels.append(level)
                    vals.append(field((o + local + 0.5) * cs))
    s = np.concatenate(vals)
    s = (s - s.min()) / max(s.max() - s.min(), 1e-30)
    B = len(origins)
    offsets = np.arange(B, dtype=np.uint64) * np.uint64(block ** 3)
    return dict(origins=np.array(origins, np.int32), dims=np.full((B, 3), block, np.int32),
                levels=np.array(levels, np.int32), offsets=offsets, refinements=np.array([2, 2], np.int32),
                scalars=s.astype(np.float32), extent=2 * n)
