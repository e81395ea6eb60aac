"""Benchmark: structured-volume delta tracking on one MI355X.

Workload (the volume configuration the headline metric is quoted on): a
512^3 fp32 field (32 seeded Gaussian blobs + 4-octave value noise), 256-entry
RGBA ramp transfer function, 16^3-voxel macrocells, delta tracking with
single scattering, 1920x1080, 16 paths per pixel. One step = one full
frame (all 16 paths) through bnamd_volume_render.

Prints ONE JSON line (value = volume Msamples/s), with the roofline of the
dominant kernel (its HBM traffic from the committed rocprofv3 byte-counter
summary), the CPU oracle timed on a bounded sample of the same frame, and
the other workloads as extra figures: surface megakernel Mrays/s (1,024
tessellated spheres), the multi-view megakernel vs wavefront renderer
(heightfield + spheres + cylinders, 64 views), the tetrahedral volume with
its iso-surface, frame readback (tiled accumulation to linear float4 /
sRGB RGBA8), and the emission-absorption forward/adjoint pass.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
RAY_IO_BYTES = 48  # per traced ray: 32 B in (org, dir, tmin, tmax) + 16 B hit out


def surface_roofline(rays, node_visits, prim_bytes, ms):
    """Algorithmic bytes of surface rays: 48 B ray I/O + 128 B per BVH8 node
    visit + the bytes of each primitive test (36 triangle, 16 sphere, 28
    cylinder), counted by the kernels; as GB/s and fraction of HBM peak."""
    b = RAY_IO_BYTES * rays + 128 * node_visits + prim_bytes
    gbps = b / (ms * 1e-3) / 1e9
    return {"alg_bytes": int(b), "achieved_GBps": round(gbps, 1), "frac": round(gbps / HBM_PEAK_GBPS, 4),
            "node_visits_per_ray": round(node_visits / max(rays, 1), 2),
            "prim_bytes_per_ray": round(prim_bytes / max(rays, 1), 1)}


def umesh_roofline(counted, ms):
    """Algorithmic bytes of unstructured-mesh sampling: 128 B per BVH node
    visited by point location + the element tests' bytes (80 B per tet: four
    indices, positions and scalars; more for larger cells), counted by the
    instrumented kernel over one frame of the same paths; as GB/s over the
    uncounted kernel's frame time and fraction of HBM peak."""
    b = 128 * counted["nodeVisits"] + counted["cellBytes"]
    gbps = b / (ms * 1e-3) / 1e9
    return {"alg_bytes": int(b), "achieved_GBps": round(gbps, 1), "frac": round(gbps / HBM_PEAK_GBPS, 4),
            "node_visits": int(counted["nodeVisits"]), "cell_tests": int(counted["cellTests"]),
            "node_visits_per_sample": round(counted["nodeVisits"] / max(counted["samples"], 1), 2),
            "cell_tests_per_sample": round(counted["cellTests"] / max(counted["samples"], 1), 2)}


def volume_roofline(samples, dda_steps, ms, value_bytes=4):
    """Volume algorithmic bytes (eight corner values per field sample, 32 B
    at fp32; 4 B per macrocell step) as GB/s and fraction of HBM peak."""
    b = 8 * value_bytes * samples + 4 * dda_steps
    gbps = b / (ms * 1e-3) / 1e9
    return {"alg_bytes": int(b), "achieved_GBps": round(gbps, 1), "frac": round(gbps / HBM_PEAK_GBPS, 4)}


METRIC = "Mrays/s + volume Msamples/s on 1 MI355X; adjoint ms/iter vs forward"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


_W2 = {}


def workload2(n):
    """Workload 2's field, camera and transfer function (made once)."""
    from barney_amd import scenes
    if n not in _W2:
        t0 = time.time()
        _W2[n] = scenes.workload2(n=n)
        log(f"field {n}^3 generated in {time.time() - t0:.1f}s")
    return _W2[n]


def volume_bench(api, steps, warmup, n=512, cpu=True, cpu_seconds=15.0):
    w2 = workload2(n)
    vol = api.Volume(w2["field"], cells_per_mc=w2["cells_per_mc"])
    vol.set_tf(w2["tf"], (0.0, 1.0), w2["base_density"])
    W, H = w2["size"]
    spp = w2["spp"]
    cam = api.camera(**w2["camera"])
    acc = api.DeviceArray.empty((H, W, 4), np.float32)
    for k in range(warmup):
        vol.render(cam, api.render_params(W, H, spp=spp, accum_id=0), accum=acc)
    api.capi.check(api.capi.load().bnamd_synchronize())
    samples = dda = rays = zero = 0
    kernel_ms = 0.0
    t0 = time.perf_counter()
    for k in range(steps):
        o = vol.render(cam, api.render_params(W, H, spp=spp, accum_id=0), accum=acc)
        samples += o["stats"]["samples"]
        dda += o["stats"]["ddaSteps"]
        zero += o["stats"]["zeroMajorantSteps"]
        rays += o["stats"]["raysTraced"]
        kernel_ms += o["stats"]["ms"]
        log(f"volume step {k}: {o['stats']['ms']:.1f} ms, {o['stats']['samples'] / 1e6:.1f} M samples")
    api.capi.check(api.capi.load().bnamd_synchronize())
    wall = time.perf_counter() - t0
    img = acc.numpy()
    res = {"samples": samples, "ddaSteps": dda, "zeroSteps": zero, "rays": rays, "wall_s": wall,
           "kernel_ms": kernel_ms,
           "steps": steps, "image_mean": float(img[..., :3].mean() / spp)}
    if cpu:
        base, oimg = volume_cpu_baseline(w2, cpu_seconds)
        y0, y1 = base.pop("_rows")
        res["cpu"] = base
        g = img[y0:y1]
        # the full-size frame against the oracle on the band it rendered
        res["vs_oracle"] = {"rows": [y0, y1 - 1],
                            "pixels_bit_exact": float(np.all(g == oimg, axis=-1).mean()),
                            "max_abs_diff": float(np.abs(g - oimg).max())}
        log(f"volume vs oracle: {res['vs_oracle']}")
    return res


def cpu_threads():
    threads = os.cpu_count() or 1
    return min(threads, int(os.environ.get("OMP_NUM_THREADS", threads)))


def band_timer(render_rows, H, seconds, unit, what, keep_rows=False):
    """Times the oracle over a band of rows around the frame's middle, sized
    from a one-row calibration to take about 'seconds'. render_rows(y0, y1)
    returns the count of work units (rays or samples) it did. keep_rows adds
    the band's row range as "_rows" (for the caller to pop)."""
    mid = H // 2
    t0 = time.perf_counter()
    render_rows(mid, mid + 1)
    dt = time.perf_counter() - t0
    rows = int(max(1, min(H, seconds / max(dt, 1e-3))))
    y0 = max(0, mid - rows // 2)
    units, reps = 0, 0
    t0 = time.perf_counter()
    while True:  # the band again until the time is used (fast workloads)
        units += render_rows(y0, y0 + rows)
        reps += 1
        dt = time.perf_counter() - t0
        if dt >= 0.5 * seconds or reps >= 1000:
            break
    res = {"value": units / dt / 1e6, "unit": unit, "cores": cpu_threads(), "kind": "port",
           "sample": f"rows {y0}..{y0 + rows - 1} of {what}, {reps}x ({units} {unit.split('/')[0][1:]}, "
                     f"{dt:.1f}s)"}
    if keep_rows:
        res["_rows"] = (y0, y0 + rows)
    return res


def volume_cpu_baseline(w2, seconds):
    """The oracle on host cores over a band of rows of the same frame;
    returns the baseline and the oracle's image of the band."""
    from oracle import refcpu
    rv = refcpu.Volume(w2["field"], cells_per_mc=w2["cells_per_mc"])
    rv.set_tf(w2["tf"], (0.0, 1.0), w2["base_density"])
    W, H = w2["size"]
    cam = refcpu.camera(**w2["camera"])
    p = refcpu.render_params(W, H, spp=w2["spp"])
    acc = np.zeros((H, W, 4), np.float32)
    base = band_timer(lambda y0, y1: rv.render(cam, p, y0, y1, cpu_threads(), accum=acc)[2], H, seconds,
                      "Msamples/s", f"the {W}x{H}x{w2['spp']}spp frame", keep_rows=True)
    y0, y1 = base["_rows"]
    return base, acc[y0:y1]


def surface_cpu_baseline(w1, seconds):
    """Workload 1 in the oracle (its own binned-SAH BVH over the same
    triangles): rays traced per second over a band of rows."""
    from oracle import refcpu
    ref = refcpu.Scene([dict(kind=0, vertices=v, indices=t, color=c, user_id=k)
                        for k, (v, t, c) in enumerate(w1["geoms"])])
    W, H = w1["size"]
    cam = refcpu.camera(**w1["camera"])
    p = refcpu.render_params(W, H, spp=1)
    acc = np.zeros((H, W, 4), np.float32)
    base = band_timer(lambda y0, y1: ref.render(cam, p, y0, y1, cpu_threads(), accum=acc)[2], H, seconds,
                      "Mrays/s", f"the {W}x{H} frame, 1 spp", keep_rows=True)
    return base, acc[base["_rows"][0]:base["_rows"][1]]


def _oracle_material(m):
    if m["type"] == "matte":
        return dict(type="matte", color=m["color"])
    if m["type"] == "pbr":
        return dict(type="pbr", color=m["color"], metallic=m.get("metallic", 0.0), roughness=m["roughness"],
                    transmission=0.0, ior=1.45)
    return dict(type="pbr", color=(1.0, 1.0, 1.0), metallic=0.0, roughness=0.5, transmission=1.0, ior=m["ior"])


def batch_cpu_baseline(w3, seconds):
    """Workload 3 in the oracle: the first view's rows, 4 paths per pixel,
    same materials, depth and Russian roulette; all rays traced per second.
    Returns the baseline and the oracle's image of those rows."""
    from oracle import refcpu
    t0 = time.time()
    ref = refcpu.Scene([dict(kind=g["kind"], vertices=g["vertices"], indices=g.get("indices"),
                             radii=g.get("radii"), material=_oracle_material(w3["materials"][g["material"]]))
                        for g in w3["geoms"]])
    log(f"workload 3 oracle scene built in {time.time() - t0:.1f}s")
    W, H = w3["widths"][0], w3["heights"][0]
    cam = refcpu.camera(**w3["cameras"][0])
    p = refcpu.render_params(W, H, spp=w3["spp"], ambient=1.0, bg=(-1, -1, -1, -1), max_depth=w3["max_depth"],
                             max_diffuse=w3["max_diffuse"], russian_roulette=w3["rr_depth"])
    acc = np.zeros((H, W, 4), np.float32)
    base = band_timer(lambda y0, y1: ref.render(cam, p, y0, y1, cpu_threads(), accum=acc)[2], H, seconds,
                      "Mrays/s", f"view 0 ({W}x{H}, {w3['spp']} spp)", keep_rows=True)
    y0, y1 = base["_rows"]
    return base, acc[y0:y1], ref


def band_vs_oracle(gpu_img, base, oimg):
    """Per-pixel agreement of a full-size GPU frame with the oracle's band:
    the fraction of pixels equal bit for bit, and within 1e-3 relative."""
    y0, y1 = base.pop("_rows")
    g = gpu_img[y0:y1]
    rel = np.abs(g - oimg) / np.maximum(np.abs(oimg), 1e-3)
    return {"rows": [y0, y1 - 1], "pixels_bit_exact": float(np.all(g == oimg, axis=-1).mean()),
            "pixels_within_1e-3": float((rel.max(-1) <= 1e-3).mean())}


def image_vs_oracle(g, r):
    """Error of a band of a GPU image against the oracle's (RGB, the
    accumulated 4-path sums): RMSE relative to the oracle's mean, and the
    relative difference of the band means (the converged error: per-path RNG
    streams re-seed from hit distances, so a last-ulp libm difference changes
    the rest of a path, not its distribution)."""
    g = g[..., :3].astype(np.float64)
    r = r[..., :3].astype(np.float64)
    m = r.mean()
    return {"rel_rmse": float(np.sqrt(((g - r) ** 2).mean()) / m), "mean_rel_diff": float((g.mean() - m) / m)}


def umesh_cpu_baseline(w5, seconds):
    """Workload 5 in the oracle (its own tet BVH and rasterTet macrocells):
    field samples per second over a band of rows, volume + iso-surface."""
    from oracle import refcpu
    t0 = time.time()
    u = refcpu.UMesh(w5["vertices"], w5["scalars"], w5["tets"])
    u.set_tf(w5["tf"], (0.0, 1.0), w5["base_density"])
    log(f"workload 5 oracle mesh built in {time.time() - t0:.1f}s")
    W, H = w5["size"]
    cam = refcpu.camera(**w5["camera"])
    p = refcpu.render_params(W, H, spp=w5["spp"], ambient=1.0, bg=(-1, -1, -1, -1))
    acc = np.zeros((H, W, 4), np.float32)
    base = band_timer(lambda y0, y1: u.render(cam, p, volume=True, iso=w5["iso"], iso_color=w5["iso_color"], y0=y0,
                                              y1=y1, threads=cpu_threads(), accum=acc)[2], H, seconds,
                      "Msamples/s", f"the {W}x{H}x{w5['spp']}spp frame", keep_rows=True)
    return base, acc[base["_rows"][0]:base["_rows"][1]]


def surface_bench(api, steps=10, warmup=2, cpu=True, cpu_seconds=10.0):
    from barney_amd import scenes
    w1 = scenes.workload1()
    geoms, mats = [], []
    for k, (v, t, c) in enumerate(w1["geoms"]):
        geoms.append(api.Geom(api.GEOM_TRIANGLES, v, t, material=k, user_id=k))
        mats.append(api.matte(c))
    world = api.World([geoms], materials=mats)
    info = world.info()
    W, H = w1["size"]
    cam = api.camera(**w1["camera"])
    acc = api.DeviceArray.empty((H, W, 4), np.float32)
    for _ in range(warmup):
        api.render_surface(world, cam, api.render_params(W, H, spp=1), accum=acc)
    ms = []
    for _ in range(steps):  # every frame traces the same rays
        o = api.render_surface(world, cam, api.render_params(W, H, spp=1), accum=acc)
        rays, nodes, pbytes = o["stats"]["raysTraced"], o["stats"]["nodeVisits"], o["stats"]["primBytes"]
        ms.append(o["stats"]["ms"])
    t = float(np.median(ms))
    res = {"Mrays_per_s": rays / (t * 1e-3) / 1e6, "rays_per_frame": rays,
           "ms_per_frame": t, "frames": steps, "bvh_build_ms": info["buildMs"], "prims": info["numPrims"],
           "bvh_nodes": info["numNodes"], "bvh_depth": info["depth"],
           "roofline": surface_roofline(rays, nodes, pbytes, t)}
    if cpu:
        base, oimg = surface_cpu_baseline(w1, cpu_seconds)
        res["vs_oracle"] = band_vs_oracle(acc.numpy(), base, oimg)
        res["cpu_baseline"] = base
    return res


def lit_mixed_bench(api, steps=10, warmup=2, spp=4, cpu=True, cpu_seconds=8.0):
    """Workload 1's spheres lit by a quad, a directional and two point lights
    next to the ambient environment, with a 128^3 blob-and-noise structured
    volume among them (a model holding geometry and a volume):
    bnamd_render_mixed at 1024x768, 4 paths per pixel."""
    from barney_amd import scenes
    from oracle import refcpu
    w1 = scenes.workload1()
    geoms, mats = [], []
    for k, (v, t, c) in enumerate(w1["geoms"]):
        geoms.append(api.Geom(api.GEOM_TRIANGLES, v, t, material=k, user_id=k))
        mats.append(api.matte(c))
    world = api.World([geoms], materials=mats)
    quads = [[20.0, 110.0, 20.0, 60.0, 0.0, 0.0, 0.0, 0.0, 60.0, 3.0, 3.0, 2.5]]
    dirs = [[0.3, -1.0, 0.4, 1.0, 0.95, 0.9, 1.5]]
    points = [[50.0, 60.0, 50.0, 1.0, 0.8, 0.6, 4.0e4, float("nan")], [10.0, 30.0, 80.0, 0.5, 0.7, 1.0, 1.0, 300.0]]
    n = 128
    f = scenes.structured_field(n=n, blobs=12, sigma=(4.0, 16.0), octaves=3, seed=0x5EED0906)
    vkw = dict(origin=(30.0, 20.0, 30.0), spacing=(40.0 / (n - 1),) * 3, cells_per_mc=8)
    vol = api.Volume(f, **vkw)
    vol.set_tf(scenes.ramp_tf(), (0.0, 1.0), 0.3)
    lights = api.Lights(quads, dirs, points)
    W, H = w1["size"]
    cam = api.camera(**w1["camera"])
    acc = api.DeviceArray.empty((H, W, 4), np.float32)
    params = api.render_params(W, H, spp=spp, ambient=0.5, lights=lights)
    for _ in range(warmup):
        api.render_mixed(world, vol, cam, params, accum=acc)
    ms = []
    for _ in range(steps):  # every frame traces the same paths
        o = api.render_mixed(world, vol, cam, params, accum=acc)
        ms.append(o["stats"]["ms"])
    t = float(np.median(ms))
    rays, samples = o["stats"]["raysTraced"], o["stats"]["samples"]
    res = {"ms_per_frame": t, "frames": steps, "spp": spp, "rays_per_frame": rays,
           "Mrays_per_s": rays / (t * 1e-3) / 1e6, "volume_samples_per_frame": samples,
           "lights": {"quad": 1, "directional": 1, "point": 2}, "volume": n}
    if cpu:
        ref = refcpu.Scene([dict(kind=0, vertices=v, indices=t, color=c, user_id=k)
                            for k, (v, t, c) in enumerate(w1["geoms"])])
        rv = refcpu.Volume(f, **vkw)
        rv.set_tf(scenes.ramp_tf(), (0.0, 1.0), 0.3)
        rl = refcpu.Lights(quads, dirs, points)
        rp = refcpu.render_params(W, H, spp=spp, ambient=0.5, lights=rl)
        racc = np.zeros((H, W, 4), np.float32)
        rcam = refcpu.camera(**w1["camera"])
        base = band_timer(lambda y0, y1: ref.render_mixed(rv, rcam, rp, y0, y1, cpu_threads(), accum=racc)[2], H,
                          cpu_seconds, "Mrays/s", f"the {W}x{H} frame, {spp} spp", keep_rows=True)
        y0, y1 = base["_rows"]
        res["vs_oracle"] = band_vs_oracle(acc.numpy(), base, racc[y0:y1])
        res["cpu_baseline"] = base
    return res


def batch_bench(api, steps=10, cpu=True, cpu_seconds=10.0):
    """Workload 3: 10M-triangle heightfield + 200k spheres + 100k cylinders,
    60/30/10 diffuse/glossy/dielectric, depth <= 8 with Russian roulette,
    64 views of uneven size, 4 paths per pixel: megakernel vs wavefront."""
    from barney_amd import scenes
    t0 = time.time()
    w3 = scenes.workload3()
    geoms = [api.Geom(g["kind"], g["vertices"], g.get("indices"), g.get("radii"), material=g["material"])
             for g in w3["geoms"]]
    world = api.World([geoms], materials=scenes.product_materials(api, w3["materials"]))
    info = world.info()
    log(f"workload 3 built in {time.time() - t0:.1f}s (BVH {info['buildMs']:.0f} ms, {info['numPrims']} prims)")
    views = api.Views([api.camera(**c) for c in w3["cameras"]], w3["widths"], w3["heights"])
    p = api.render_params(0, 0, spp=w3["spp"], ambient=1.0, bg=(-1, -1, -1, -1), max_depth=w3["max_depth"],
                          max_diffuse=w3["max_diffuse"], russian_roulette=w3["rr_depth"])
    acc = api.DeviceArray.empty((views.num_pixels, 4), np.float32)
    res = {}
    import tempfile
    with tempfile.TemporaryDirectory() as td:
        api.batch_autotune(world, views, p, os.path.join(td, "tune.json"))
        with open(os.path.join(td, "tune.json")) as f:
            res["autotune"] = json.load(f)
    for name, mode in (("megakernel", api.BATCH_MEGAKERNEL), ("wavefront", api.BATCH_WAVEFRONT)):
        api.render_batch(world, views, p, mode, accum=acc)  # warm-up (graph capture)
        ms, rays = [], 0
        for _ in range(steps):
            o = api.render_batch(world, views, p, mode, accum=acc)
            ms.append(o["stats"]["ms"])
            rays = o["stats"]["raysTraced"]
            nodes, pbytes = o["stats"]["nodeVisits"], o["stats"]["primBytes"]
        img = acc.numpy()
        res[name] = {"view0": views.image(img, 0), "ms_per_frame": float(np.median(ms)), "rays_per_frame": int(rays),
                     "Mrays_per_s": rays / (np.median(ms) * 1e-3) / 1e6, "image": img,
                     "roofline": surface_roofline(rays, nodes, pbytes, float(np.median(ms)))}
        log(f"workload 3 {name}: {np.median(ms):.1f} ms, {rays / 1e6:.1f} M rays")
    view0 = {k: res[k].pop("view0") for k in ("megakernel", "wavefront")}
    a, b = res["megakernel"].pop("image"), res["wavefront"].pop("image")
    res["max_abs_diff_wavefront_vs_megakernel"] = float(np.abs(a - b).max())
    res["wavefront_speedup"] = res["wavefront"]["Mrays_per_s"] / res["megakernel"]["Mrays_per_s"]
    res["pixels"] = views.num_pixels
    res["bvh_build_ms"] = info["buildMs"]
    if cpu:
        base, oimg, ref = batch_cpu_baseline(w3, cpu_seconds)
        y0, y1 = base.pop("_rows")
        res["cpu_baseline"] = base
        res["vs_oracle"] = {"rows": [y0, y1 - 1], **{k: image_vs_oracle(view0[k][y0:y1], oimg)
                                                     for k in ("megakernel", "wavefront")}}
        log(f"workload 3 vs oracle (view 0 rows {y0}..{y1 - 1}): {res['vs_oracle']}")
        res["converged_vs_oracle"] = batch_converged(api, w3, world, ref)
        log(f"workload 3 converged (256 paths per pixel) vs oracle: {res['converged_vs_oracle']}")
    return res


def batch_converged(api, w3, world, ref, spp=256, rows=16):
    """The workload's target measure: view 0 rendered with 256 paths per
    pixel by both GPU renderers (view 0 alone in the batch) and, over a band
    of rows, by the oracle; the RMSE of the per-pixel means against the
    oracle's, relative to the oracle's band mean (the same per-path streams
    on both sides)."""
    from oracle import refcpu
    W, H = w3["widths"][0], w3["heights"][0]
    y0 = max(0, H // 2 - rows // 2)
    y1 = min(H, y0 + rows)
    c = w3["cameras"][0]
    views0 = api.Views([api.camera(**c)], [W], [H])
    kw = dict(ambient=1.0, bg=(-1, -1, -1, -1), max_depth=w3["max_depth"], max_diffuse=w3["max_diffuse"],
              russian_roulette=w3["rr_depth"])
    racc = np.zeros((H, W, 4), np.float32)
    t0 = time.time()
    ref.render(refcpu.camera(**c), refcpu.render_params(W, H, spp=spp, **kw), y0, y1, cpu_threads(), accum=racc)
    r = racc[y0:y1, :, :3].astype(np.float64) / spp
    out = {"spp": spp, "view": 0, "rows": [y0, y1 - 1], "oracle_s": round(time.time() - t0, 2)}
    for name, mode in (("megakernel", api.BATCH_MEGAKERNEL), ("wavefront", api.BATCH_WAVEFRONT)):
        acc = api.DeviceArray.empty((W * H, 4), np.float32)
        o = api.render_batch(world, views0, api.render_params(0, 0, spp=spp, **kw), mode, accum=acc)
        g = views0.image(acc.numpy(), 0)[y0:y1, :, :3].astype(np.float64) / spp
        out[name] = {"rel_rmse": float(np.sqrt(((g - r) ** 2).mean()) / r.mean()),
                     "mean_rel_diff": float((g.mean() - r.mean()) / r.mean()), "ms": o["stats"]["ms"]}
    return out


def umesh_bench(api, steps=10, cpu=True, cpu_seconds=10.0):
    """Workload 5: 128^3 jittered hex grid split into 5 tets per hex (10.5M
    tets), per-vertex Gaussian-blob scalars; BVH point location with
    barycentric sampling, rasterTet macrocell majorants, delta tracking plus
    an iso-surface at 0.5; 1024x1024, 8 paths per pixel."""
    from barney_amd import scenes
    t0 = time.time()
    w5 = scenes.workload5()
    m = api.UMesh(w5["vertices"], w5["scalars"], w5["tets"])
    m.set_tf(w5["tf"], (0.0, 1.0), w5["base_density"])
    info = m.info()
    log(f"workload 5 built in {time.time() - t0:.1f}s (BVH {info['buildMs']:.0f} ms, MC {info['mcMs']:.0f} ms)")
    W, H = w5["size"]
    cam = api.camera(**w5["camera"])
    p = api.render_params(W, H, spp=w5["spp"], ambient=1.0, bg=(-1, -1, -1, -1))
    acc = api.DeviceArray.empty((H, W, 4), np.float32)
    counted = m.render(cam, p, volume=True, iso=w5["iso"], iso_color=w5["iso_color"], accum=acc, count=True)["stats"]
    m.render(cam, p, volume=True, iso=w5["iso"], iso_color=w5["iso_color"], accum=acc)
    ms, st = [], None
    for _ in range(steps):
        o = m.render(cam, p, volume=True, iso=w5["iso"], iso_color=w5["iso_color"], accum=acc)
        st = o["stats"]
        ms.append(st["ms"])
    t = float(np.median(ms))
    assert counted["samples"] == st["samples"], "the counted frame took other paths"
    log(f"workload 5: {t:.1f} ms/frame, {st['samples'] / 1e6:.1f} M field samples")
    res = {"ms_per_frame": t, "field_samples_per_frame": int(st["samples"]),
            "Msamples_per_s": st["samples"] / (t * 1e-3) / 1e6, "rays_per_frame": int(st["raysTraced"]),
            "Mrays_per_s": st["raysTraced"] / (t * 1e-3) / 1e6, "dda_steps_per_frame": int(st["ddaSteps"]),
            "tets": info["numTets"], "bvh_nodes": info["numNodes"], "bvh_depth": info["bvhDepth"],
            "mc_dims": list(info["mcDims"]), "bvh_build_ms": info["buildMs"], "mc_build_ms": info["mcMs"],
            "image_mean": float(acc.numpy()[..., :3].mean() / w5["spp"]),
            "roofline": umesh_roofline(counted, t), "counted_frame_ms": counted["ms"],
            "cpu_baseline": None}
    if cpu:
        base, oimg = umesh_cpu_baseline(w5, cpu_seconds)
        res["vs_oracle"] = band_vs_oracle(acc.numpy(), base, oimg)
        res["cpu_baseline"] = base
    return res


def umesh_cells_bench(api, n=40, steps=10, cpu=True, cpu_seconds=8.0, polyhedra=False):
    """General unstructured cells (an extra workload beside the five; the same path as
    workload 5 with pyramids, prisms and hexahedra next to tets): the sheared
    mixed mesh of scenes.mixed_cells on an n^3 lattice, delta tracking plus
    an iso-surface at 0.5, 512x512, 4 paths per pixel. polyhedra=True writes
    half of the hexes and half of the prisms as VTK face-stream polyhedra."""
    from barney_amd import scenes
    from oracle import refcpu
    t0 = time.time()
    V, S, cells = scenes.mixed_cells(n, polyhedra=polyhedra)
    m = api.UMesh(V, S, cells=cells)
    tf = scenes.ramp_tf()
    m.set_tf(tf, (0.0, 1.0), 0.3)
    info = m.info()
    log(f"mixed cells built in {time.time() - t0:.1f}s ({info['numTets']} cells)")
    W, H, spp = 512, 512, 4
    c = n / 2.0
    pos = (c - 1.1 * n, c + 0.8 * n, c - 1.7 * n)
    camargs = dict(position=pos, direction=(c - pos[0], c - pos[1], c - pos[2]), fovy=40.0)
    p = api.render_params(W, H, spp=spp, ambient=1.0, bg=(-1, -1, -1, -1))
    acc = api.DeviceArray.empty((H, W, 4), np.float32)
    cam = api.camera(**camargs)
    counted = m.render(cam, p, volume=True, iso=0.5, accum=acc, count=True)["stats"]
    m.render(cam, p, volume=True, iso=0.5, accum=acc)
    ms, st = [], None
    for _ in range(steps):
        o = m.render(cam, p, volume=True, iso=0.5, accum=acc)
        st = o["stats"]
        ms.append(st["ms"])
    t = float(np.median(ms))
    assert counted["samples"] == st["samples"], "the counted frame took other paths"
    log(f"mixed cells: {t:.1f} ms/frame, {st['samples'] / 1e6:.1f} M field samples")
    res = {"cells": info["numTets"], "types": {int(k): int(v) for k, v in zip(*np.unique(cells[2], return_counts=True))},
           "ms_per_frame": t, "field_samples_per_frame": int(st["samples"]),
           "Msamples_per_s": st["samples"] / (t * 1e-3) / 1e6, "rays_per_frame": int(st["raysTraced"]),
           "roofline": umesh_roofline(counted, t)}
    if cpu:
        r = refcpu.UMesh(V, S, cells=cells)
        r.set_tf(tf, (0.0, 1.0), 0.3)
        rp = refcpu.render_params(W, H, spp=spp, ambient=1.0, bg=(-1, -1, -1, -1))
        rcam = refcpu.camera(**camargs)
        racc = np.zeros((H, W, 4), np.float32)
        res["cpu_baseline"] = band_timer(
            lambda y0, y1: r.render(rcam, rp, volume=True, iso=0.5, y0=y0, y1=y1, threads=cpu_threads(),
                                    accum=racc)[2], H, cpu_seconds, "Msamples/s", f"the {W}x{H}x{spp}spp frame")
    return res


def amr_bench(api, n=128, block=16, steps=10, cpu=True, cpu_seconds=8.0):
    """Block-structured AMR (an extra workload beside the five; the structured-volume
    path tracer with BlockStructuredField's sampler): scenes.amr_field with a
    128^3-cell coarse level and a 128^3-cell fine level in 16^3 blocks
    (1,024 blocks, 4.2 M cells), delta tracking, 1024x1024, 4 paths per
    pixel."""
    from barney_amd import scenes
    from oracle import refcpu
    t0 = time.time()
    f = scenes.amr_field(n, block)
    args = (f["origins"], f["dims"], f["levels"], f["offsets"], f["refinements"], f["scalars"])
    a = api.Amr(*args)
    tf = scenes.ramp_tf()
    a.set_tf(tf, (0.0, 1.0), 0.2)
    info = a.info()
    log(f"amr built in {time.time() - t0:.1f}s ({info['numBlocks']} blocks, build {info['buildMs']:.0f} ms)")
    W, H, spp = 1024, 1024, 4
    c = f["extent"] / 2.0
    pos = (c - 1.3 * f["extent"], c + 0.7 * f["extent"], c - 1.6 * f["extent"])
    camargs = dict(position=pos, direction=(c - pos[0], c - pos[1], c - pos[2]), fovy=35.0)
    p = api.render_params(W, H, spp=spp, ambient=1.0, bg=(-1, -1, -1, -1))
    acc = api.DeviceArray.empty((H, W, 4), np.float32)
    cam = api.camera(**camargs)
    a.render(cam, p, accum=acc)
    ms, st = [], None
    for _ in range(steps):
        o = a.render(cam, p, accum=acc)
        st = o["stats"]
        ms.append(st["ms"])
    t = float(np.median(ms))
    log(f"amr: {t:.1f} ms/frame, {st['samples'] / 1e6:.1f} M samples")
    res = {"blocks": info["numBlocks"], "cells": int(len(f["scalars"])), "ms_per_frame": t,
           "samples_per_frame": int(st["samples"]), "Msamples_per_s": st["samples"] / (t * 1e-3) / 1e6,
           "dda_steps_per_frame": int(st["ddaSteps"]), "rays_per_frame": int(st["raysTraced"]),
           "roofline": volume_roofline(st["samples"], st["ddaSteps"], t)}
    if cpu:
        r = refcpu.Amr(*args)
        r.set_tf(tf, (0.0, 1.0), 0.2)
        rp = refcpu.render_params(W, H, spp=spp, ambient=1.0, bg=(-1, -1, -1, -1))
        rcam = refcpu.camera(**camargs)
        racc = np.zeros((H, W, 4), np.float32)
        res["cpu_baseline"] = band_timer(lambda y0, y1: r.render(rcam, rp, y0, y1, cpu_threads(), accum=racc)[2], H,
                                         cpu_seconds, "Msamples/s", f"the {W}x{H}x{spp}spp frame")
    return res


def nanovdb_bench(api, value_type, n=512, steps=10, cpu=True, cpu_seconds=6.0):
    """NanoVDB field (an extra workload beside the five): workload 2's 512^3
    field written as a NanoVDB grid of 'value_type' values (one voxel per
    scalar, unit voxels), its camera, transfer function and frame
    (1920x1080, 16 paths per pixel), rendered by the volume path tracer with
    barney's NanoVDB sampler; the oracle renders a band of the same frame."""
    from barney_amd import nanovdb
    from oracle import refcpu
    w2 = workload2(n)
    t0 = time.time()
    grid = nanovdb.from_dense(w2["field"], value_type=value_type)
    g = api.NanoVDB(grid)
    g.set_tf(w2["tf"], (0.0, 1.0), w2["base_density"])
    info = g.info()
    log(f"nanovdb {value_type}: {grid.size / 1e6:.0f} MB grid written and loaded in {time.time() - t0:.1f}s "
        f"(macrocells {info['mcDims']}, build {info['buildMs']:.0f} ms)")
    W, H = w2["size"]
    spp = w2["spp"]
    p = api.render_params(W, H, spp=spp)
    cam = api.camera(**w2["camera"])
    acc = api.DeviceArray.empty((H, W, 4), np.float32)
    g.render(cam, p, accum=acc)
    ms, st = [], None
    for _ in range(steps):
        o = g.render(cam, p, accum=acc)
        st = o["stats"]
        ms.append(st["ms"])
    t = float(np.median(ms))
    vb = {"float": 4, "fp16": 2, "fp8": 1, "fp4": 0.5}.get(value_type, 4)
    log(f"nanovdb {value_type}: {t:.2f} ms/frame, {st['samples'] / 1e6:.1f} M samples")
    res = {"value_type": value_type, "grid_bytes": int(grid.size), "ms_per_frame": t,
           "samples_per_frame": int(st["samples"]), "Msamples_per_s": st["samples"] / (t * 1e-3) / 1e6,
           "dda_steps_per_frame": int(st["ddaSteps"]), "rays_per_frame": int(st["raysTraced"]),
           "build_ms": info["buildMs"], "roofline": volume_roofline(st["samples"], st["ddaSteps"], t, vb)}
    if cpu:
        r = refcpu.NanoVDB(grid)
        r.set_tf(w2["tf"], (0.0, 1.0), w2["base_density"])
        rp = refcpu.render_params(W, H, spp=spp)
        rcam = refcpu.camera(**w2["camera"])
        racc = np.zeros((H, W, 4), np.float32)
        base = band_timer(lambda y0, y1: r.render(rcam, rp, y0, y1, cpu_threads(), accum=racc)[2], H,
                          cpu_seconds, "Msamples/s", f"the {W}x{H}x{spp}spp frame", keep_rows=True)
        y0, y1 = base.pop("_rows")
        res["cpu_baseline"] = base
        img = acc.numpy()[y0:y1]
        res["vs_oracle"] = {"rows": [y0, y1 - 1],
                            "pixels_bit_exact": float(np.all(img == racc[y0:y1], axis=-1).mean())}
        log(f"nanovdb {value_type} vs oracle: {res['vs_oracle']}")
    return res


def fb_bench(api, reps=20):
    """Frame readback of a 3840x2160 frame (8160 tiles of 32x32): tiled
    accumulation -> linear float4 + normals, -> sRGB RGBA8, and the resolve of
    a linear accumulation buffer to float4. Timed with HIP events on one
    stream around `reps` launches; algorithmic bytes per pixel: 56 (float4 +
    normal in and out), 20 (float4 in, RGBA8 out), 32 (float4 in and out)."""
    import ctypes as C
    import torch
    lib = api.capi.load()
    W, H = 3840, 2160
    descs = api.fb_tile_descs(W, H)
    n = descs.shape[0]
    rng = np.random.default_rng(7)
    tiles = api.DeviceArray(rng.random((n, 1024 * 7), dtype=np.float32) * 4.0)
    accum = api.DeviceArray(rng.random((H, W, 4), dtype=np.float32) * 4.0)
    color = api.DeviceArray.empty((H, W, 4), np.float32)
    normal = api.DeviceArray.empty((H, W, 3), np.float32)
    rgba8 = api.DeviceArray.empty((H, W), np.uint32)
    s = torch.cuda.Stream()
    sp = C.c_void_p(s.cuda_stream)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def timed(fn):
        fn()
        s.synchronize()
        e0.record(s)
        for _ in range(reps):
            fn()
        e1.record(s)
        e1.synchronize()
        return e0.elapsed_time(e1) / reps

    px = W * H
    legs = {
        "linearize_float4_normal": (lambda: lib.bnamd_fb_linearize(tiles.ptr, descs.ptr, n, W, H, 0.25, 0, color.ptr,
                                                                    normal.ptr, sp), 56),
        "linearize_srgb8": (lambda: lib.bnamd_fb_linearize(tiles.ptr, descs.ptr, n, W, H, 0.25, 2, rgba8.ptr, None,
                                                           sp), 20),
        "resolve_float4": (lambda: lib.bnamd_fb_resolve(accum.ptr, W, H, 0.25, 0, color.ptr, sp), 32),
        "resolve_srgb8": (lambda: lib.bnamd_fb_resolve(accum.ptr, W, H, 0.25, 2, rgba8.ptr, sp), 20),
    }
    res = {"width": W, "height": H, "tiles": n}
    for name, (fn, bpp) in legs.items():
        ms = timed(fn)
        gbps = bpp * px / (ms * 1e-3) / 1e9
        res[name] = {"us": round(ms * 1e3, 2), "GBps_alg": round(gbps, 1), "frac": round(gbps / HBM_PEAK_GBPS, 3)}
        log(f"readback {name}: {ms * 1e3:.1f} us, {gbps:.0f} GB/s")
    api.capi.check(lib.bnamd_synchronize())
    return res


# the sources a kernel's code is built from: a traffic figure measured on
# other sources is not this kernel's
KERNEL_SOURCES = {"k_render_volume": ["volume.hip", "volume_render.h", "volume.h", "shade.h", "envmap.h", "fmath.h",
                                      "common.h", "world.h", "bvh8.h", "runtime.h", "amr.h", "lights.h"]}


def sources_sha256(kernel):
    import hashlib
    h = hashlib.sha256()
    for name in KERNEL_SOURCES[kernel]:
        with open(os.path.join(ROOT, "barney_amd", "csrc", name), "rb") as f:
            h.update(name.encode() + b"\0" + f.read())
    return h.hexdigest()


def profile_traffic(kernel):
    """HBM bytes per launch of 'kernel' from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (profiles/traffic.json), each counter
    scaled by what it reads for a copy kernel of known size; None when the
    figures were measured on other kernel sources than the ones built."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)[kernel]
    except (OSError, KeyError, ValueError):
        return None
    if d.get("sources_sha256") != sources_sha256(kernel):
        log(f"profiles/traffic.json: {kernel} was measured on other sources; traffic left null")
        return None
    return int((d["fetch_kib"] * d["fetch_scale"] + d["write_kib"] * d["write_scale"]) * 1024)


def ea_bench(steps=10):
    """Workload 4: 256^3 density, 64 orbit views at 512^2, L2 loss against
    images of a second transfer function; forward and checkpointed adjoint."""
    import torch
    from barney_amd import ea, scenes
    w4 = scenes.workload4()
    n = w4["density"].shape[0]
    W, H = w4["size"]
    setup = ea.EASetup((n, n, n), w4["cameras"], W, H, step=w4["step"], base_density=w4["base_density"],
                       checkpoint_every=int(os.environ.get("BNAMD_EA_CHECKPOINT", "8")))
    dens = torch.tensor(w4["density"], device="cuda", requires_grad=True)
    tf = torch.tensor(w4["tf_init"], device="cuda", requires_grad=True)
    with torch.no_grad():
        target = ea.ea_render(dens.detach(), torch.tensor(w4["tf_target"], device="cuda"), setup)
    fwd, bwd = [], []
    for k in range(steps + 1):
        dens.grad = None
        tf.grad = None
        img = ea.ea_render(dens, tf, setup)
        loss = ((img - target) ** 2).sum()
        loss.backward()
        torch.cuda.synchronize()
        if k > 0:
            fwd.append(setup.last_forward["ms"])
            bwd.append(setup.last_backward["ms"])
    f, b = float(np.median(fwd)), float(np.median(bwd))
    samples = setup.last_forward["samples"]
    res = {"forward_ms": f, "adjoint_ms": b, "adjoint_over_forward": b / f, "samples_per_iter": samples,
           "forward_Msamples_per_s": samples / (f * 1e-3) / 1e6,
           "forward_GBps_alg": 32.0 * samples / (f * 1e-3) / 1e9,
           "adjoint_GBps_alg": 96.0 * samples / (b * 1e-3) / 1e9,
           "loss": float(loss.detach()), "grad_tf_norm": float(tf.grad.norm()),
           "grad_density_norm": float(dens.grad.norm())}
    res["gradient_check"] = ea_directional_check(ea, setup, dens, tf, target)
    log(f"workload 4 gradient check: {res['gradient_check']}")
    return res


def ea_directional_check(ea, setup, dens, tf, target, eps=1e-3):
    """The full-size adjoint against central differences along a seeded
    direction v of the transfer function, then of the density grid:
    <grad L, v> from the adjoint vs (L(x + eps v) - L(x - eps v)) / (2 eps),
    the loss summed in float64. v_i = sign(g_i) u_i with u uniform in
    [0.5, 1] and g the adjoint's gradient: a direction of random signs makes
    <grad L, v> a sum of millions of cancelling terms, far below the
    terms' own scale, so any rounding in the difference quotient swamps
    it; with the gradient's signs the terms add, and an entry whose sign
    the adjoint got wrong still counts against the difference quotient.
    The density direction is zero on voxels within 2 eps of the transfer
    function's domain ends: the lookup clamps there (most of the grid is
    empty space at exactly 0), so L has a kink that a central difference
    halves while the adjoint takes one side."""
    import torch
    gen = torch.Generator(device="cuda").manual_seed(11)

    def loss(d, t):
        with torch.no_grad():
            return float(((ea.ea_render(d, t, setup) - target).double() ** 2).sum())

    out = {}
    for name, x in (("tf", tf), ("density", dens)):
        u = torch.rand(x.shape, generator=gen, device="cuda") * 0.5 + 0.5
        v = torch.sign(x.grad) * u
        if name == "density":
            xd = x.detach()
            v = torch.where((xd > 2 * eps) & (xd < 1 - 2 * eps), v, torch.zeros_like(v))
        dot = float((x.grad.double() * v.double()).sum())
        xp, xm = (x.detach() + eps * v), (x.detach() - eps * v)
        if name == "tf":
            fd = (loss(dens.detach(), xp) - loss(dens.detach(), xm)) / (2 * eps)
        else:
            fd = (loss(xp, tf.detach()) - loss(xm, tf.detach())) / (2 * eps)
        out[name] = {"adjoint": dot, "central_difference": fd, "rel_diff": abs(fd - dot) / abs(fd), "eps": eps}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-surface", action="store_true")
    ap.add_argument("--no-ea", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-umesh", action="store_true")
    ap.add_argument("--no-fb", action="store_true")
    ap.add_argument("--no-nanovdb", action="store_true")
    ap.add_argument("--n", type=int, default=512, help="volume resolution (default: the workload's 512)")
    ap.add_argument("--copy-calib", action="store_true",
                    help="also run a 1 GiB streaming copy kernel (HBM byte-counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the path shards by image tiles / views)", file=sys.stderr)
        sys.exit(2)
    if not (args.no_ea and args.no_fb):
        # torch must bring up the HIP runtime before libbnamd does when both
        # share the process
        import torch  # noqa: F401
        torch.cuda.init()
    from barney_amd import api
    api.init(0)
    if args.copy_calib:
        src = api.DeviceArray.empty((1 << 28,), np.float32).zero()
        dst = api.DeviceArray.empty((1 << 28,), np.float32)
        for _ in range(3):
            api.capi.check(api.capi.load().bnamd_copy_kernel(dst.addr, src.addr, 1 << 30, None))
        log("copy calibration: 3 x 1 GiB copies done")
    vb = volume_bench(api, args.steps, args.warmup, n=args.n, cpu=not args.no_cpu)
    per_launch_samples = vb["samples"] / vb["steps"]
    per_launch_dda = vb["ddaSteps"] / vb["steps"]
    kernel_s = vb["kernel_ms"] * 1e-3 / vb["steps"]
    alg_bytes = 32.0 * per_launch_samples + 4.0 * per_launch_dda
    achieved = alg_bytes / kernel_s / 1e9
    value = vb["samples"] / vb["wall_s"] / 1e6
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(vb["wall_s"] * 1e3 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "structured volume 512^3 fp32, 32 Gaussian blobs + 4-octave value noise, "
                               "256-entry RGBA TF, 16^3-voxel macrocells, delta tracking + single scattering, "
                               "1920x1080, 16 spp",
                   "width": 1920, "height": 1080, "spp": 16, "volume": args.n, "cells_per_mc": 16},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": profile_traffic("k_render_volume"),
                     "kernel": "k_render_volume", "avg_launch_ms": round(kernel_s * 1e3, 3),
                     "alg_bytes_per_launch": int(alg_bytes)},
        "cpu_baseline": vb.get("cpu"),
        "extra": {"samples_per_step": int(per_launch_samples), "dda_steps_per_step": int(per_launch_dda),
                  "zero_majorant_steps_per_step": vb["zeroSteps"] // vb["steps"],
                  "rays_per_step": vb["rays"] // vb["steps"], "image_mean": vb["image_mean"],
                  "vs_oracle": vb.get("vs_oracle")},
    }
    if not args.no_surface:
        out["extra"]["surface_workload1"] = surface_bench(api, cpu=not args.no_cpu)
        out["extra"]["lit_mixed_workload1"] = lit_mixed_bench(api, cpu=not args.no_cpu)
    if not args.no_batch:
        out["extra"]["batch_workload3"] = batch_bench(api, cpu=not args.no_cpu)
    if not args.no_umesh:
        out["extra"]["umesh_workload5"] = umesh_bench(api, cpu=not args.no_cpu)
        out["extra"]["umesh_mixed_cells"] = umesh_cells_bench(api, cpu=not args.no_cpu)
        out["extra"]["umesh_mixed_polyhedra"] = umesh_cells_bench(api, cpu=not args.no_cpu, cpu_seconds=4.0,
                                                                  polyhedra=True)
        out["extra"]["amr_two_levels"] = amr_bench(api, cpu=not args.no_cpu)
    if not args.no_nanovdb:
        out["extra"]["nanovdb_float"] = nanovdb_bench(api, "float", n=args.n, cpu=not args.no_cpu)
        out["extra"]["nanovdb_fp8"] = nanovdb_bench(api, "fp8", n=args.n, cpu=not args.no_cpu)
    if not args.no_fb:
        out["extra"]["readback_3840x2160"] = fb_bench(api)
    if not args.no_ea:
        out["extra"]["ea_workload4"] = ea_bench()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
