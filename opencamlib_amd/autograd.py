"""torch.autograd.Function of the drop-cutter (the differentiable path).

DropCutter.apply(cl_xy, vertices, radius, mesh, spec, z0, precision) returns
the dropped heights z (n,) of CL points cl_xy (n, 2) on the device mesh
`mesh` (a capi.DeviceMesh built from `vertices`).  Backward runs the HIP
backward kernel (ocl_hip_drop_backward) on the active contacts the forward
recorded and returns dL/d(cl_xy), dL/d(vertices) (ntri, 9) and dL/d(radius).

vertices and radius are inputs only so that autograd can route their
gradients: the forward reads the mesh handle and the cutter spec, so a
caller that changes the vertices must rebuild the DeviceMesh, and radius
must equal the spec's radius (d / 2; checked, ValueError otherwise;
composites have no single radius and take radius None).  radius may be a
float or a tensor; gradients are returned only for the inputs that require
them.  All tensors are float64 on the HIP device; both passes run on torch's
current stream.
"""
from __future__ import annotations

import torch

from . import capi


def _ptr(t):
    return t.data_ptr() if t is not None else 0


class DropCutter(torch.autograd.Function):
    @staticmethod
    def forward(ctx, cl_xy, vertices, radius, mesh, spec, z0=-1e3, precision="mixed"):
        if cl_xy.device.type != "cuda" or cl_xy.dtype != torch.float64:
            raise ValueError("cl_xy must be a float64 tensor on the HIP device")
        if spec["kind"] == "composite":
            if radius is not None:
                raise ValueError("a composite cutter has no single radius: pass radius=None")
        else:
            r = float(radius)
            if abs(r - spec["d"] / 2.0) > 1e-12 * (1.0 + abs(r)):
                raise ValueError(f"radius {r} differs from the cutter spec's radius {spec['d'] / 2.0}")
        n = cl_xy.shape[0]
        dev = cl_xy.device
        cl = torch.empty((n, 3), dtype=torch.float64, device=dev)
        cl[:, :2] = cl_xy
        cl[:, 2] = z0
        z = torch.empty(n, dtype=torch.float64, device=dev)
        cc = torch.empty((n, 3), dtype=torch.float64, device=dev)
        ty = torch.empty(n, dtype=torch.uint8, device=dev)
        tri = torch.empty(n, dtype=torch.int32, device=dev)
        ft = torch.empty(n, dtype=torch.uint8, device=dev)
        stream = torch.cuda.current_stream(dev).cuda_stream
        mesh.drop_device(spec, cl.data_ptr(), n, precision, z.data_ptr(), cc.data_ptr(), ty.data_ptr(),
                         tri.data_ptr(), ft.data_ptr(), stream=stream)
        ctx.mesh, ctx.spec, ctx.precision = mesh, spec, precision
        ctx.ntri = vertices.shape[0]
        ctx.save_for_backward(cl, cc, ty, tri, ft)
        ctx.mark_non_differentiable(cc)
        return z

    @staticmethod
    def backward(ctx, gz):
        cl, cc, ty, tri, ft = ctx.saved_tensors
        n = cl.shape[0]
        dev = cl.device
        want_cl, want_tri, want_r = ctx.needs_input_grad[0], ctx.needs_input_grad[1], ctx.needs_input_grad[2]
        gz = gz.contiguous().to(torch.float64)
        dz = torch.empty((n, 3), dtype=torch.float64, device=dev) if want_cl else None
        gtri = torch.zeros((ctx.ntri, 9), dtype=torch.float64, device=dev) if want_tri else None
        gr = torch.zeros(1, dtype=torch.float64, device=dev) if want_r else None
        if want_cl or want_tri or want_r:
            stream = torch.cuda.current_stream(dev).cuda_stream
            ctx.mesh.drop_backward(ctx.spec, cl.data_ptr(), n, cc.data_ptr(), ty.data_ptr(), tri.data_ptr(),
                                   ft.data_ptr(), d_grad_z=gz.data_ptr(), precision=ctx.precision,
                                   d_dz=_ptr(dz), d_grad_tri=_ptr(gtri), d_grad_r=_ptr(gr), stream=stream)
        g_cl = gz[:, None] * dz[:, :2] if want_cl else None
        return g_cl, gtri, gr.reshape(()) if want_r else None, None, None, None, None


def drop_z(cl_xy, vertices, radius, mesh, spec, z0=-1e3, precision="mixed"):
    """Differentiable dropped heights (see DropCutter)."""
    return DropCutter.apply(cl_xy, vertices, radius, mesh, spec, z0, precision)
