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
