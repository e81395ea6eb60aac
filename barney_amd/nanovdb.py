"""NanoVDB grid buffers from dense arrays (host-side data preparation).

barney's "NanoVDB" scalar field takes a NanoVDB grid buffer as its data
(barney/volume/NanoVDB.cu:190-235, the buffer normally written by the NanoVDB
tools). This module writes such buffers from a dense fp32 array in the
NanoVDB 32.x memory layout (3rdParty/nanovdb/include/nanovdb/NanoVDB.h):

    GridData (672 B) | TreeData (64 B) | RootData (64 B) + tiles (32 B each)
    | upper internal nodes (32^3, 270400 B) | lower internal nodes (16^3,
    33856 B) | leaves (8^3)

with float leaves (2144 B) or leaves quantized per leaf to Fp4 / Fp8 / Fp16
codes or FpN codes of one chosen bit width (value = code * quantum + min, the
codes rounded as CreateNanoGrid does without dithering: uint(encode * (v -
min) + 0.5)). Voxel (i, j, k) of a (nz, ny, nx) array (x fastest) sits at
index (i0 + i, j0 + j, k0 + k); the map is a uniform scale plus translation,
world = voxel_size * index + origin.

Every voxel of the array is active; voxels of a leaf outside the array hold
the background value and are inactive.
"""
import numpy as np

GRID_TYPES = {"float": 1, "fp4": 13, "fp8": 14, "fp16": 15, "fpn": 16}
_MAGIC_NUMB = 0x304244566F6E614E  # "NanoVDB0"
_MAGIC_GRID = 0x314244566F6E614E  # "NanoVDB1"
_VERSION = (32 << 21) | (7 << 10) | 0
_UPPER_SIZE, _LOWER_SIZE = 270400, 33856
_FIXED_LEAF_SIZE = {"float": 2144, "fp4": 352, "fp8": 608, "fp16": 1120}


def _leaf_size(value_type, bits):
    if value_type == "fpn":
        return 96 + bits * 64
    return _FIXED_LEAF_SIZE[value_type]


def _encode_leaves(vals, value_type, bits):
    """vals: [L, 512] fp32 leaf values (all voxels). Returns (mins, quanta,
    codes bytes [L, n]) as CreateNanoGrid's processLeafs computes them."""
    width = {"fp4": 4, "fp8": 8, "fp16": 16}.get(value_type, bits)
    mn = vals.min(axis=1)
    mx = vals.max(axis=1)
    units = float((1 << width) - 1)
    quantum = ((mx - mn) / np.float32(units)).astype(np.float32)
    with np.errstate(divide="ignore", invalid="ignore"):
        if width >= 16:
            enc = units / (mx.astype(np.float64) - mn.astype(np.float64))
            q = enc[:, None] * (vals - mn[:, None]).astype(np.float64) + 0.5
        else:
            enc = (np.float32(units) / (mx - mn)).astype(np.float32)
            q = (enc[:, None] * (vals - mn[:, None]) + np.float32(0.5)).astype(np.float32)
        q = np.where(np.isfinite(q), q, 0.0)
    codes = np.clip(q, 0, units).astype(np.uint32)
    L = len(vals)
    if width == 16:
        out = codes.astype(np.uint16).view(np.uint8).reshape(L, 1024)
    elif width == 8:
        out = codes.astype(np.uint8)
    else:  # pack 8 / width codes per byte, low bits first
        per = 8 // width
        c = codes.reshape(L, 512 // per, per)
        b = np.zeros((L, 512 // per), np.uint32)
        for j in range(per):
            b |= c[:, :, j] << (j * width)
        out = b.astype(np.uint8)
    return mn.astype(np.float32), quantum, out


def from_dense(values, voxel_size=1.0, origin=(0.0, 0.0, 0.0), index_origin=(0, 0, 0), value_type="float",
               fpn_bits=8, background=0.0, name="density"):
    """A NanoVDB grid buffer (np.uint8 array) holding 'values' (nz, ny, nx)."""
    if value_type not in GRID_TYPES:
        raise ValueError(f"value_type must be one of {sorted(GRID_TYPES)}")
    if value_type == "fpn" and fpn_bits not in (1, 2, 4, 8, 16):
        raise ValueError("fpn_bits must be 1, 2, 4, 8 or 16")
    if not voxel_size > 0:
        raise ValueError("voxel_size must be positive")
    v = np.ascontiguousarray(values, np.float32)
    if v.ndim != 3 or min(v.shape) < 1:
        raise ValueError("values must be a non-empty (nz, ny, nx) array")
    nz, ny, nx = v.shape
    i0 = np.asarray(index_origin, np.int64)
    lo = i0
    hi = i0 + np.array([nx, ny, nz]) - 1
    bg = np.float32(background)

    # leaves: the 8^3 blocks that the array touches, padded with background
    l_lo, l_hi = lo >> 3, hi >> 3
    LX, LY, LZ = (l_hi - l_lo + 1).tolist()
    pad = np.full((LZ * 8, LY * 8, LX * 8), bg, np.float32)
    act = np.zeros(pad.shape, bool)
    o = lo - (l_lo << 3)
    pad[o[2]:o[2] + nz, o[1]:o[1] + ny, o[0]:o[0] + nx] = v
    act[o[2]:o[2] + nz, o[1]:o[1] + ny, o[0]:o[0] + nx] = True
    # [LZ, 8z, LY, 8y, LX, 8x] -> leaf (lx, ly, lz), voxel offset x << 6 | y << 3 | z
    lv = pad.reshape(LZ, 8, LY, 8, LX, 8).transpose(4, 2, 0, 5, 3, 1).reshape(-1, 512)
    la = act.reshape(LZ, 8, LY, 8, LX, 8).transpose(4, 2, 0, 5, 3, 1).reshape(-1, 512)
    gx, gy, gz = np.meshgrid(np.arange(LX), np.arange(LY), np.arange(LZ), indexing="ij")
    leaf_org = (np.stack([gx.ravel(), gy.ravel(), gz.ravel()], 1) + l_lo) << 3  # [L, 3]

    # parents: lower nodes (128^3), upper nodes (4096^3)
    low_org_of_leaf = leaf_org & ~np.int64(127)
    up_org_of_leaf = leaf_org & ~np.int64(4095)
    lows, low_idx = np.unique(low_org_of_leaf, axis=0, return_inverse=True)
    ups, up_idx_of_low = np.unique(lows & ~np.int64(4095), axis=0, return_inverse=True)
    low_idx = low_idx.ravel()
    up_idx_of_low = up_idx_of_low.ravel()

    def local(org, parent, shift, log2):
        rel = (org - parent) >> shift
        return (rel[:, 0] << (2 * log2)) | (rel[:, 1] << log2) | rel[:, 2]

    leaf_slot = local(leaf_org, low_org_of_leaf, 3, 4)  # offset in its lower node
    low_slot = local(lows, lows & ~np.int64(4095), 7, 5)  # offset in its upper node
    # breadth-first order: uppers as in ups; lowers grouped by upper then slot;
    # leaves grouped by lower (in lower order) then slot
    low_order = np.lexsort((low_slot, up_idx_of_low))
    low_rank = np.empty_like(low_order)
    low_rank[low_order] = np.arange(len(low_order))
    leaf_order = np.lexsort((leaf_slot, low_rank[low_idx]))
    del up_org_of_leaf

    nU, nL, nF = len(ups), len(lows), len(leaf_org)
    lsize = _leaf_size(value_type, fpn_bits)
    root_off = 672 + 64
    tiles_off = root_off + 64
    up_off = tiles_off + 32 * nU
    low_off = up_off + _UPPER_SIZE * nU
    leaf_off = low_off + _LOWER_SIZE * nL
    total = leaf_off + lsize * nF
    buf = np.zeros(total, np.uint8)

    def put(off, arr):
        a = np.ascontiguousarray(arr).view(np.uint8).ravel()
        buf[off:off + a.size] = a

    active_vals = v.ravel()
    vmin, vmax = np.float32(active_vals.min()), np.float32(active_vals.max())
    vavg = np.float32(active_vals.astype(np.float64).mean())
    vdev = np.float32(active_vals.astype(np.float64).std())

    # ---- leaves
    lv_o, la_o, org_o = lv[leaf_order], la[leaf_order], leaf_org[leaf_order]
    leaf_pos = np.empty(nF, np.int64)  # leaf index -> byte offset
    leaf_pos[leaf_order] = leaf_off + lsize * np.arange(nF, dtype=np.int64)
    head = np.zeros((nF, 96), np.uint8)
    head_i = head.view(np.int32)  # [nF, 24]
    head_i[:, 0:3] = org_o.astype(np.int32)
    # mBBoxDif (3 B) and mFlags (1 B) at byte 12..15: the active voxels' box
    bits_arr = la_o.reshape(nF, 8, 8, 8)
    anyx = bits_arr.any(axis=(2, 3))
    anyy = bits_arr.any(axis=(1, 3))
    anyz = bits_arr.any(axis=(1, 2))
    first = lambda m: m.argmax(axis=1)  # noqa: E731
    last = lambda m: 7 - m[:, ::-1].argmax(axis=1)  # noqa: E731
    bmin = np.stack([first(anyx), first(anyy), first(anyz)], 1)
    bmax = np.stack([last(anyx), last(anyy), last(anyz)], 1)
    head_i[:, 0:3] = (org_o + bmin).astype(np.int32)
    head[:, 12:15] = (bmax - bmin).astype(np.uint8)
    log_bits = {1: 0, 2: 1, 4: 2, 8: 3, 16: 4}[fpn_bits]
    head[:, 15] = 2 | 16 | ((log_bits << 5) if value_type == "fpn" else 0)
    # value mask: bit n of word n >> 6
    mask_bits = np.packbits(la_o.reshape(nF, 512), axis=1, bitorder="little")  # [nF, 64] bytes
    head[:, 16:80] = mask_bits
    if value_type == "float":
        act_v = np.where(la_o, lv_o, np.nan)
        st = np.zeros((nF, 4), np.float32)
        with np.errstate(invalid="ignore"):
            st[:, 0] = np.nanmin(act_v, axis=1)
            st[:, 1] = np.nanmax(act_v, axis=1)
            st[:, 2] = np.nanmean(act_v, axis=1)
            st[:, 3] = np.nanstd(act_v, axis=1)
        head[:, 80:96] = st.view(np.uint8)
        body = lv_o.view(np.uint8).reshape(nF, 2048)
    else:
        mn, quantum, codes = _encode_leaves(lv_o, value_type, fpn_bits)
        head[:, 80:84] = mn.view(np.uint8).reshape(nF, 4)
        head[:, 84:88] = quantum.view(np.uint8).reshape(nF, 4)
        body = codes
    leaves = np.zeros((nF, lsize), np.uint8)
    leaves[:, :96] = head
    leaves[:, 96:96 + body.shape[1]] = body
    buf[leaf_off:leaf_off + nF * lsize] = leaves.ravel()

    # ---- lower nodes
    low_pos = np.empty(nL, np.int64)
    low_pos[low_order] = low_off + _LOWER_SIZE * np.arange(nL, dtype=np.int64)
    by_low = np.argsort(low_idx, kind="stable")
    bounds = np.searchsorted(low_idx[by_low], np.arange(nL + 1))
    for li in range(nL):
        base = int(low_pos[li])
        sel = by_low[bounds[li]:bounds[li + 1]]
        node = np.zeros(_LOWER_SIZE, np.uint8)
        ni = node[:32].view(np.int32)
        ni[0:3] = lows[li].astype(np.int32)
        ni[3:6] = (lows[li] + 127).astype(np.int32)
        node[24:32] = np.frombuffer(np.uint64(2 | 4).tobytes(), np.uint8)
        child = np.zeros(4096, bool)
        child[leaf_slot[sel]] = True
        node[544:1056] = np.packbits(child, bitorder="little")
        table = np.zeros(4096, np.int64)
        table_f = table.view(np.float32).reshape(4096, 2)
        table_f[:, 0] = bg
        table[leaf_slot[sel]] = leaf_pos[sel] - base
        node[1056:1072] = np.array([vmin, vmax, vavg, vdev], np.float32).view(np.uint8)
        node[1088:] = table.view(np.uint8)
        buf[base:base + _LOWER_SIZE] = node

    # ---- upper nodes and root tiles
    up_pos = [up_off + _UPPER_SIZE * u for u in range(nU)]
    for u in range(nU):
        base = up_pos[u]
        sel = np.nonzero(up_idx_of_low == u)[0]
        node = np.zeros(_UPPER_SIZE, np.uint8)
        ni = node[:32].view(np.int32)
        ni[0:3] = ups[u].astype(np.int32)
        ni[3:6] = (ups[u] + 4095).astype(np.int32)
        node[24:32] = np.frombuffer(np.uint64(2 | 4).tobytes(), np.uint8)
        child = np.zeros(32768, bool)
        child[low_slot[sel]] = True
        node[4128:8224] = np.packbits(child, bitorder="little")
        table = np.zeros(32768, np.int64)
        table.view(np.float32).reshape(32768, 2)[:, 0] = bg
        table[low_slot[sel]] = low_pos[sel] - base
        node[8224:8240] = np.array([vmin, vmax, vavg, vdev], np.float32).view(np.uint8)
        node[8256:] = table.view(np.uint8)
        buf[base:base + _UPPER_SIZE] = node

    def key(o):
        x, y, z = (int(c) & 0xFFFFFFFF for c in o)
        return (z >> 12) | ((y >> 12) << 21) | ((x >> 12) << 42)

    keys = [key(o) for o in ups]
    for r, u in enumerate(sorted(range(nU), key=lambda j: keys[j])):
        t = np.zeros(32, np.uint8)
        t[0:8] = np.frombuffer(np.uint64(keys[u]).tobytes(), np.uint8)
        t[8:16] = np.frombuffer(np.int64(up_pos[u] - root_off).tobytes(), np.uint8)
        t[20:24] = np.frombuffer(bg.tobytes(), np.uint8)
        buf[tiles_off + 32 * r:tiles_off + 32 * r + 32] = t

    root = np.zeros(64, np.uint8)
    root[0:24] = np.concatenate([lo, hi]).astype(np.int32).view(np.uint8)
    root[24:28] = np.frombuffer(np.uint32(nU).tobytes(), np.uint8)
    root[28:48] = np.array([bg, vmin, vmax, vavg, vdev], np.float32).view(np.uint8)
    buf[root_off:root_off + 64] = root

    # ---- tree
    tree = np.zeros(64, np.uint8)
    tree[0:32] = np.array([leaf_off - 672, low_off - 672, up_off - 672, root_off - 672], np.int64).view(np.uint8)
    tree[32:44] = np.array([nF, nL, nU], np.uint32).view(np.uint8)
    tree[56:64] = np.frombuffer(np.uint64(nx * ny * nz).tobytes(), np.uint8)
    buf[672:736] = tree

    # ---- grid header
    s = float(voxel_size)
    t = np.asarray(origin, np.float64)
    g = np.zeros(672, np.uint8)
    put_g = lambda off, a: g.__setitem__(slice(off, off + a.nbytes), a.view(np.uint8).ravel())  # noqa: E731
    put_g(0, np.array([_MAGIC_NUMB, 0xFFFFFFFFFFFFFFFF], np.uint64))
    put_g(16, np.array([_VERSION, 2 | 4 | 32, 0, 1], np.uint32))
    put_g(32, np.array([total], np.uint64))
    nm = name.encode()[:255]
    g[40:40 + len(nm)] = np.frombuffer(nm, np.uint8)
    inv = 1.0 / s
    eye = np.eye(3)
    m = 296
    put_g(m, (eye * s).astype(np.float32).ravel())
    put_g(m + 36, (eye * inv).astype(np.float32).ravel())
    put_g(m + 72, np.array([*t, 1.0], np.float32))
    put_g(m + 88, (eye * s).ravel().astype(np.float64))
    put_g(m + 160, (eye * inv).ravel().astype(np.float64))
    put_g(m + 232, np.array([*t, 1.0], np.float64))
    wlo = s * lo.astype(np.float64) + t
    whi = s * (hi + 1).astype(np.float64) + t
    put_g(560, np.concatenate([wlo, whi]).astype(np.float64))
    put_g(608, np.array([s, s, s], np.float64))
    put_g(632, np.array([2, GRID_TYPES[value_type]], np.uint32))
    put_g(640, np.array([total], np.int64))
    put_g(664, np.array([_MAGIC_GRID], np.uint64))
    buf[0:672] = g
    return buf
