"""Writes the NanoVDB golden vectors (tests/golden/nanovdb_<case>.npz).

Each case is produced by oracle/_ref/nvdb_ref, which is built (oracle/Makefile)
from oracle/nvdb_ref.cpp and the NanoVDB 32.7.0 headers that the reference
vendors: the grid buffer comes from nanovdb::tools::createNanoGrid, the samples
from barney's NanoVDBDataSampler statement over nanovdb::math::createSampler<1>
(barney/volume/NanoVDB.h:98-127), the macrocell ranges from NanoVDBData's
computeMCs statement over the library's ReadAccessor (NanoVDB.cu:28-102).

    python tests/golden/make_nanovdb_vectors.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
GEN = os.path.join(ROOT, "oracle", "_ref", "nvdb_ref")

# case name -> (value type, seed, random points, extra argument)
CASES = {
    "float": ("float", 7, 2000, None),
    "fp4": ("fp4", 7, 2000, None),
    "fp8": ("fp8", 7, 2000, None),
    "fp16": ("fp16", 7, 2000, None),
    "fpn": ("fpn", 7, 2000, None),
    "tiles": ("float", 11, 2000, "tiles"),
}


def parse_meta(path):
    out = {}
    with open(path) as f:
        for line in f:
            key, *vals = line.split()
            conv = []
            for v in vals:
                if "p" in v or "x" in v:
                    conv.append(float.fromhex(v))
                elif "." in v or "e" in v:
                    conv.append(float(v))
                else:
                    conv.append(int(v))
            out[key] = conv
    return out


def make(name):
    vtype, seed, npts, extra = CASES[name]
    with tempfile.TemporaryDirectory() as d:
        cmd = [GEN, d, vtype, str(seed), str(npts)] + ([extra] if extra else [])
        subprocess.run(cmd, check=True)
        meta = parse_meta(os.path.join(d, "meta.txt"))
        grid = np.fromfile(os.path.join(d, "grid.nvdb"), np.uint8)
        points = np.fromfile(os.path.join(d, "points.f32"), np.float32).reshape(-1, 3)
        samples = np.fromfile(os.path.join(d, "samples.f32"), np.float32)
        ranges = np.fromfile(os.path.join(d, "ranges.f32"), np.float32).reshape(-1, 2)
    np.savez_compressed(
        os.path.join(HERE, f"nanovdb_{name}.npz"), grid=grid, points=points, samples=samples, ranges=ranges,
        grid_type=np.int32(meta["gridType"][0]), index_lo=np.int32(meta["indexLo"]), index_hi=np.int32(meta["indexHi"]),
        world_lo=np.float32(meta["worldLo"]), world_hi=np.float32(meta["worldHi"]), mc_dims=np.int32(meta["mcDims"]),
        mc_spacing=np.float32(meta["mcSpacing"]), voxel_size=np.float64(meta["voxelSize"][0]),
        nodes=np.int32(meta["nodes"]), root_tiles=np.int32(meta["rootTiles"][0]))
    print(f"nanovdb_{name}.npz: {grid.size} grid bytes, {len(points)} points, {len(ranges)} macrocells")


if __name__ == "__main__":
    if not os.path.exists(GEN):
        sys.exit(f"{GEN} is missing: build it with 'make -C oracle _ref/nvdb_ref'")
    for name in (sys.argv[1:] or CASES):
        make(name)
