"""Differentiable emission-absorption rendering on the GPU.

EARender is a torch.autograd.Function over bnamd_ea_forward /
bnamd_ea_backward (include/bnamd.h): the forward stores sparse transmittance
checkpoints, the backward replays each checkpoint segment in reverse and
returns gradients for the density grid and the transfer function. Tensors
must live on the GPU; the kernels run on torch's current stream.

Import torch (or this module) before the first barney_amd.api call in a
process: torch has to initialise the HIP runtime that both then share.
"""
import ctypes as C

import torch

from . import capi
from .capi import check


class bnamd_ea_volume(C.Structure):
    _fields_ = [("density", C.c_void_p), ("dims", C.c_int * 3), ("gridOrigin", C.c_float * 3),
                ("gridSpacing", C.c_float * 3), ("tf", C.c_void_p), ("numTF", C.c_int), ("domainLo", C.c_float),
                ("domainHi", C.c_float), ("baseDensity", C.c_float), ("step", C.c_float)]


class bnamd_ea_views(C.Structure):
    _fields_ = [("width", C.c_int), ("height", C.c_int), ("numViews", C.c_int),
                ("cameras", C.POINTER(capi.bnamd_camera)), ("checkpointEvery", C.c_int)]


class bnamd_ea_stats(C.Structure):
    _fields_ = [("samples", C.c_longlong), ("ms", C.c_float)]


class EASetup:
    """Static description of an EA render: grid geometry, TF domain, step,
    views. The density and TF tensors are the differentiable inputs."""

    def __init__(self, dims, cameras, width, height, step=0.5, origin=(0, 0, 0), spacing=(1, 1, 1),
                 domain=(0.0, 1.0), base_density=1.0, checkpoint_every=16):
        from . import api
        api.init()
        self.dims = tuple(int(d) for d in dims)  # (nx, ny, nz)
        self.width, self.height = int(width), int(height)
        self.cams = (capi.bnamd_camera * len(cameras))()
        for k, c in enumerate(cameras):
            self.cams[k] = api.camera(c["position"], c["direction"], c.get("up", (0, 1, 0)), c["fovy"])
        self.num_views = len(cameras)
        self.step, self.origin, self.spacing = float(step), origin, spacing
        self.domain, self.base_density = domain, float(base_density)
        self.checkpoint_every = int(checkpoint_every)
        self.last_forward = None
        self.last_backward = None

    def structs(self, density, tf):
        v = bnamd_ea_volume()
        v.density = density.data_ptr()
        v.dims[:] = list(self.dims)
        v.gridOrigin[:] = [float(x) for x in self.origin]
        v.gridSpacing[:] = [float(x) for x in self.spacing]
        v.tf = tf.data_ptr()
        v.numTF = tf.shape[0]
        v.domainLo, v.domainHi = float(self.domain[0]), float(self.domain[1])
        v.baseDensity = self.base_density
        v.step = self.step
        w = bnamd_ea_views()
        w.width, w.height, w.numViews = self.width, self.height, self.num_views
        w.cameras = C.cast(self.cams, C.POINTER(capi.bnamd_camera))
        w.checkpointEvery = self.checkpoint_every
        return v, w

    def checkpoint_count(self, density, tf):
        v, w = self.structs(density, tf)
        n = C.c_int(0)
        check(capi.load().bnamd_ea_checkpoint_count(C.byref(v), C.byref(w), C.byref(n)))
        return n.value


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _check_tensor(t, name):
    if not t.is_cuda or t.dtype != torch.float32 or not t.is_contiguous():
        raise ValueError(f"{name} must be a contiguous float32 CUDA tensor")


class EARender(torch.autograd.Function):
    @staticmethod
    def forward(ctx, density, tf, setup):
        _check_tensor(density, "density")
        _check_tensor(tf, "tf")
        nseg = setup.checkpoint_count(density, tf)
        rays = setup.num_views * setup.height * setup.width
        img = torch.empty((setup.num_views, setup.height, setup.width, 3), device=density.device,
                          dtype=torch.float32)
        ckpt = torch.empty((nseg, rays), device=density.device, dtype=torch.float32)
        v, w = setup.structs(density, tf)
        st = bnamd_ea_stats()
        check(capi.load().bnamd_ea_forward(C.byref(v), C.byref(w), img.data_ptr(), ckpt.data_ptr(), _stream(),
                                           C.byref(st)))
        setup.last_forward = {"ms": st.ms, "samples": st.samples}
        ctx.setup = setup
        ctx.save_for_backward(density, tf, ckpt)
        return img

    @staticmethod
    def backward(ctx, gimg):
        density, tf, ckpt = ctx.saved_tensors
        setup = ctx.setup
        gimg = gimg.contiguous().float()
        gd = torch.zeros_like(density)
        gtf = torch.zeros_like(tf)
        v, w = setup.structs(density, tf)
        st = bnamd_ea_stats()
        check(capi.load().bnamd_ea_backward(C.byref(v), C.byref(w), gimg.data_ptr(), ckpt.data_ptr(),
                                            gd.data_ptr(), gtf.data_ptr(), _stream(), C.byref(st)))
        setup.last_backward = {"ms": st.ms, "samples": st.samples}
        return gd, gtf, None


def ea_render(density, tf, setup):
    """Images (views, H, W, 3) of the density grid through the TF."""
    return EARender.apply(density, tf, setup)
