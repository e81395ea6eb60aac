"""bnamd_umesh_render_noalloc (the tet-mesh render that allocates nothing:
its counters and iso-surface mask live in the library's fixed per-GPU
scratch array) against bnamd_umesh_render, through the C ABI: the images
are bit-identical and the counters equal, on two calls in a row, the
second with a smaller mesh and frame (the scratch keeps the first call's
larger mask and counters, which the second must not see)."""
import ctypes as C

import numpy as np
import pytest

from barney_amd import capi, scenes

pytestmark = pytest.mark.gpu


def _mesh(gpu, n, seed):
    V, S, T = scenes.tet_grid(n, seed=seed)
    m = gpu.UMesh(V, S, T)
    m.set_tf(scenes.ramp_tf(), (0.0, 1.0), 0.6)
    return m


def _render(gpu, fn, mesh, cam, params):
    acc = gpu.DeviceArray.empty((params.height, params.width, 4), np.float32)
    o = capi.bnamd_umesh_render_opts()
    o.volume = 1
    o.iso = 1
    o.isoValue = 0.5
    o.isoColor[:] = [0.9, 0.7, 0.4]
    st = capi.bnamd_volume_stats()
    gpu.check(getattr(capi.load(), fn)(mesh.handle, C.byref(cam), C.byref(params), C.byref(o), acc.addr, None,
                                       C.byref(st)))
    return acc.numpy(), (st.samples, st.ddaSteps, st.raysTraced, st.zeroMajorantSteps)


def _cam(gpu, n):
    c = n / 2.0
    pos = (c - 1.2 * n, c + 0.8 * n, c - 1.9 * n)
    return gpu.camera(pos, (c - pos[0], c - pos[1], c - pos[2]), fovy=40.0)


def test_noalloc_equals_render_twice_second_smaller(gpu):
    for n, (W, H), spp in ((10, (96, 72), 8), (5, (40, 30), 4)):
        mesh = _mesh(gpu, n, 0x5EED0404 + n)
        cam = _cam(gpu, n)
        p = gpu.render_params(W, H, spp=spp, ambient=1.0, bg=(-1, -1, -1, -1), max_diffuse=1)
        a, sa = _render(gpu, "bnamd_umesh_render_noalloc", mesh, cam, p)
        b, sb = _render(gpu, "bnamd_umesh_render", mesh, cam, p)
        print(f"n {n} {W}x{H} spp {spp}: stats {sa} vs {sb}", flush=True)
        assert np.isfinite(a).all() and a[..., 3].mean() > 0.2
        assert np.array_equal(a, b)
        assert sa == sb and sa[0] > 0
