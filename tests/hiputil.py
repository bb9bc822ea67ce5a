"""Helpers for GPU tests that drive libjet_hip.so through its C ABI."""
import ctypes
import sys

import numpy as np

from jet_amd import _native as nat


def say(msg):
    """Name the case before a GPU step (flushed, so a fault still shows it)."""
    print(msg, flush=True)
    sys.stdout.flush()


def code(dtype):
    return nat.C64 if np.dtype(dtype) == np.complex64 else nat.C128


def to_device(ctx, host):
    host = np.ascontiguousarray(host)
    buf = nat.DeviceBuffer(max(host.nbytes, 8), ctx)
    buf.upload(host)
    return buf


def from_device(buf, n, dtype):
    out = np.empty(n, dtype=dtype)
    buf.download(out)
    return out


def permute(ctx, x, shape, in_modes, out_modes, layout=nat.ROW_MAJOR):
    d_in = to_device(ctx, x)
    d_out = nat.DeviceBuffer(max(x.nbytes, 8), ctx)
    nat.check(nat.lib().jet_hip_permute(ctx.handle, code(x.dtype), layout, len(shape),
                                        nat.i32(in_modes), nat.i64(shape), nat.i32(out_modes),
                                        d_in.ptr, d_out.ptr, None))
    return from_device(d_out, x.size, x.dtype)


def plan_info(plan):
    flops = ctypes.c_double()
    m, n, k, b = (ctypes.c_int64() for _ in range(4))
    kern = ctypes.c_int()
    nat.check(nat.lib().jet_hip_plan_info(plan, ctypes.byref(flops), ctypes.byref(m),
                                          ctypes.byref(n), ctypes.byref(k), ctypes.byref(b),
                                          ctypes.byref(kern)))
    return dict(flops=flops.value, M=m.value, N=n.value, K=k.value, batch=b.value,
                kernel=nat.GEMM_KERNEL_NAMES[kern.value])


def contract(ctx, a, ma, sa, b, mb, sb, mc, precision=nat.PREC_DEFAULT, layout=nat.ROW_MAJOR,
             workspace=False):
    """C[mc] = sum A[ma] B[mb] through the plan API; returns (data, plan info)."""
    dtype = a.dtype
    ext = dict(zip(ma, sa))
    ext.update(dict(zip(mb, sb)))
    sc = [ext[m] for m in mc]
    nc = int(np.prod(sc)) if sc else 1
    plan = ctypes.c_void_p()
    nat.check(nat.lib().jet_hip_contraction_plan_create(
        ctx.handle, code(dtype), precision, layout, len(ma), nat.i32(ma), nat.i64(sa), len(mb),
        nat.i32(mb), nat.i64(sb), len(mc), nat.i32(mc), ctypes.byref(plan)))
    try:
        info = plan_info(plan)
        da, db = to_device(ctx, a), to_device(ctx, b)
        dc = nat.DeviceBuffer(max(8, nc * dtype.itemsize), ctx)
        ws_ptr, ws_bytes, ws = None, 0, None
        if workspace:
            sz = ctypes.c_size_t()
            nat.check(nat.lib().jet_hip_plan_workspace_size(plan, ctypes.byref(sz)))
            if sz.value:
                ws = nat.DeviceBuffer(sz.value, ctx)
                ws_ptr, ws_bytes = ws.ptr, sz.value
        nat.check(nat.lib().jet_hip_contract(plan, da.ptr, db.ptr, dc.ptr, ws_ptr, ws_bytes,
                                             None))
        out = from_device(dc, nc, dtype)
    finally:
        nat.lib().jet_hip_plan_destroy(plan)
    return out, info


def rand_c(rng, n, dtype):
    return (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(dtype)


def rel_err(got, want):
    got = np.asarray(got, dtype=np.complex128).reshape(-1)
    want = np.asarray(want, dtype=np.complex128).reshape(-1)
    den = np.linalg.norm(want)
    return np.linalg.norm(got - want) / (den if den > 0 else 1.0)
