"""The hot path: evaluate / integrate / inner / operate / truncate / eval_grad.

Each call runs in libqtt on the GPU (there is no host fallback).  Names and
argument meaning follow the reference:
  evaluate(f, xs, dims)          src/itensornetworkfunction.jl:96-112
  integrate(f; take_sum)         src/integration.jl:6-17
  integrate(op, f; take_sum)     src/integration.jl:20-32
  operate(op, f; cutoff, maxdim) src/elementary_operators.jl:294-319
  truncate(f; cutoff, maxdim)    src/itensornetworkfunction.jl:114-118
`evaluate` is batched: xs may be an (npts, ndims) array (or a torch CUDA tensor,
in which case the values stay on the device).
"""
from __future__ import annotations

import ctypes as C
from contextlib import contextmanager

import numpy as np

from . import _capi
from .indsnetworkmap import DigitMapError
from .network import ITensorNetworkFunction, TreeOperator


class Precision:
    F64 = 0        # float64 everywhere
    F32_SPLIT = 1  # float32 (hi, lo) operands with float-float accumulation
    F32 = 2        # plain float32


def engine_context():
    """The process-wide libqtt context (device 0)."""
    return _capi.context()


@contextmanager
def on_stream(stream_ptr):
    """Run engine work on an existing HIP stream (e.g. torch.cuda.Stream().cuda_stream);
    the stream in use before is restored afterwards."""
    ctx = engine_context()
    prev = ctx.stream
    ctx.set_stream(stream_ptr)
    try:
        yield ctx
    finally:
        ctx.set_stream(prev)


@contextmanager
def _torch_stream(ctx, tensor):
    """Launch on torch's current stream of the tensor's device, so the engine's
    reads of torch inputs and writes of torch outputs are ordered with torch's
    own work on that stream (the default stream, or the one selected by
    `with torch.cuda.stream(s)`).  Torch's default stream is the null stream,
    which the context's own stream (a blocking HIP stream) is ordered with."""
    import torch
    s = torch.cuda.current_stream(tensor.device).cuda_stream or None
    prev = ctx.stream
    if s != prev:
        ctx.set_stream(s)
    try:
        yield ctx
    finally:
        if s != prev:
            ctx.set_stream(prev)


def device_network(f: ITensorNetworkFunction) -> _capi.DeviceNetwork:
    """Upload f's cores once and keep the handle on f."""
    if f.device is None:
        f.attach(_capi.DeviceNetwork.upload(engine_context(), f.topo, f.cores))
    return f.device


def device_operator(op: TreeOperator) -> _capi.DeviceOperator:
    if op._device is None:
        op._device = _capi.DeviceOperator(engine_context(), op.topo, op.cores)
    return op._device


def _points(f, xs, dims):
    nd = f.dimension()
    a = np.asarray(xs, dtype=np.float64)
    if a.ndim == 0:
        a = a.reshape(1, 1)
        single = True
    elif a.ndim == 1:
        a = a.reshape(1, -1)
        single = True
    else:
        single = False
    if dims is not None:
        dims = list(dims)
        if sorted(dims) != list(range(1, nd + 1)):
            raise ValueError("dims must list every dimension of the function once")
        full = np.empty((a.shape[0], nd))
        for j, d in enumerate(dims):
            full[:, d - 1] = a[:, j]
        a = full
    if a.shape[1] != nd:
        raise ValueError(f"points have {a.shape[1]} coordinates, the function has {nd} dimensions")
    return np.ascontiguousarray(a), single


def evaluate(f: ITensorNetworkFunction, xs, dims=None, precision: int = Precision.F64, out=None):
    """f at one point (scalar result) or at a batch of points (array result).

    With a CUDA tensor of points the call is asynchronous: it is queued on
    torch's current stream and `out` (a CUDA tensor) is ready when that stream
    reaches it.  An invalid coordinate then yields NaN at that point and a
    DigitMapError from the next `engine_context().sync()`.  A network on a
    layout with loops (LoopyNetworkFunction) is evaluated by belief
    propagation (evaluate(f, xs; alg="bp") of the reference), host points only."""
    from .loopy import LoopyNetworkFunction, evaluate_bp
    if isinstance(f, LoopyNetworkFunction):  # a layout with loops: belief propagation
        if dims is not None:
            raise ValueError("give every coordinate of a loopy network's points")
        return evaluate_bp(f, xs)
    dn = device_network(f)
    ctx = dn.ctx
    try:
        import torch
        is_torch = isinstance(xs, torch.Tensor)
    except ImportError:  # pragma: no cover
        is_torch = False
    if is_torch:
        if not xs.is_cuda or xs.dtype != torch.float64 or xs.dim() != 2 or not xs.is_contiguous():
            raise ValueError("device points must be a contiguous float64 CUDA tensor [npts, ndims]")
        if xs.shape[1] != f.dimension():
            raise ValueError("wrong number of coordinates")
        n = xs.shape[0]
        if out is None:
            out = torch.empty(n, dtype=torch.float64, device=xs.device)
        with _torch_stream(ctx, xs):
            ctx.check(_capi.lib().qtt_eval(ctx.handle, dn.handle, C.c_void_p(xs.data_ptr()), n,
                                           C.c_void_p(out.data_ptr()), int(precision)))
        return out
    pts, single = _points(f, xs, dims)
    res = np.empty(pts.shape[0])
    rc = _capi.lib().qtt_eval_host(ctx.handle, dn.handle, _capi._dptr(pts), pts.shape[0],
                                   _capi._dptr(res), int(precision))
    if rc == _capi.QTT_ERANGE:
        raise DigitMapError("a coordinate is negative or NaN (the digit map is defined for x >= 0)")
    ctx.check(rc)
    return float(res[0]) if single else res


def tune_evaluate(f: ITensorNetworkFunction, xs, out=None):
    """Time the float64 register chain kernel's launch shapes on these device
    points (a contiguous float64 CUDA tensor [npts, ndims]) and keep the fastest
    in the engine context for later evaluate calls of the same bond class and
    point-count octave.  Returns {"points_per_lane", "workgroup",
    "table_digits", "ms"} (zeros when the network does not take that kernel);
    table_digits is the number of digit sites under each environment table."""
    import torch
    dn = device_network(f)
    ctx = dn.ctx
    if not (isinstance(xs, torch.Tensor) and xs.is_cuda and xs.dtype == torch.float64 and xs.dim() == 2
            and xs.is_contiguous() and xs.shape[1] == f.dimension()):
        raise ValueError("device points must be a contiguous float64 CUDA tensor [npts, ndims]")
    if out is None:
        out = torch.empty(xs.shape[0], dtype=torch.float64, device=xs.device)
    cfg = (C.c_int * 3)()
    ms = C.c_double(0.0)
    with _torch_stream(ctx, xs):
        ctx.check(_capi.lib().qtt_tune_eval(ctx.handle, dn.handle, C.c_void_p(xs.data_ptr()), xs.shape[0],
                                            C.c_void_p(out.data_ptr()), cfg, C.byref(ms)))
    return {"points_per_lane": cfg[0], "workgroup": cfg[1], "table_digits": cfg[2], "ms": ms.value}


def _sigma(f: ITensorNetworkFunction, take_sum: bool):
    """The uniform product state used by integrate(op, f) (src/integration.jl:30)."""
    c = 1.0 if take_sum else 1.0 / f.s.base
    # a vertex without a digit has no site index to sum over
    cores = [np.full(f.topo.core_shape(v, np.ones(f.topo.nv, dtype=np.int64)),
                     c if f.topo.sdim[v] > 1 else 1.0)
             for v in range(f.topo.nv)]
    return ITensorNetworkFunction(f.s, cores, f.topo)


def integrate(op_or_f, f: ITensorNetworkFunction | None = None, take_sum: bool = False):
    """integrate(f) or integrate(op, f)."""
    if f is None:
        f = op_or_f
        dn = device_network(f)
        out = C.c_double(0.0)
        weight = 1.0 if take_sum else 1.0 / f.s.base
        dn.ctx.check(_capi.lib().qtt_integrate(dn.ctx.handle, dn.handle, weight, C.byref(out)))
        return out.value
    return inner(_sigma(f, take_sum), f, op=op_or_f)


def inner(a: ITensorNetworkFunction, b: ITensorNetworkFunction, op: TreeOperator | None = None):
    """<a|op|b> (op = None: <a|b>)."""
    da, db = device_network(a), device_network(b)
    dop = device_operator(op) if op is not None else None
    out = C.c_double(0.0)
    da.ctx.check(_capi.lib().qtt_inner(da.ctx.handle, da.handle, dop.handle if dop else None, db.handle,
                                       C.byref(out)))
    return out.value


def _cut(cutoff):
    return 0.0 if cutoff is None else float(cutoff)


def operate(op, f: ITensorNetworkFunction, cutoff=None, maxdim=None, precision: int = Precision.F64,
            truncate: bool = True, return_error: bool = False):
    """op applied to f, truncated with (cutoff, maxdim); a list of operators is
    applied in order (src/elementary_operators.jl:305-319)."""
    if isinstance(op, (list, tuple)):
        g = f
        err = 0.0
        for o in op:
            g, e = operate(o, g, cutoff, maxdim, precision, truncate, True)
            err += e
        return (g, err) if return_error else g
    dn = device_network(f)
    dop = device_operator(op)
    h = C.c_void_p()
    err = C.c_double(0.0)
    cut = _cut(cutoff) if truncate else -1.0
    dn.ctx.check(_capi.lib().qtt_apply(dn.ctx.handle, dop.handle, dn.handle, cut, int(maxdim or 0),
                                       int(precision), C.byref(h), C.byref(err)))
    g = ITensorNetworkFunction(f.s, None, f.topo, device=_capi.DeviceNetwork(dn.ctx, f.topo, h))
    return (g, err.value) if return_error else g


def truncate(f: ITensorNetworkFunction, cutoff=None, maxdim=None, return_error: bool = False):
    """A truncated copy of f (the input is left unchanged)."""
    dn = device_network(f).copy()
    err = C.c_double(0.0)
    dn.ctx.check(_capi.lib().qtt_truncate(dn.ctx.handle, dn.handle, _cut(cutoff), int(maxdim or 0),
                                          C.byref(err)))
    g = ITensorNetworkFunction(f.s, None, f.topo, device=dn)
    return (g, err.value) if return_error else g


def eval_grad(f: ITensorNetworkFunction, xs, ybar, precision: int = Precision.F64):
    """Cotangent-weighted gradient sum_m ybar[m] d f(x_m) / d cores (list of arrays)."""
    dn = device_network(f)
    pts, _ = _points(f, xs, None)
    yb = np.ascontiguousarray(np.broadcast_to(np.asarray(ybar, dtype=np.float64), (pts.shape[0],)))
    b = dn.bonds()
    grads = [np.empty(f.topo.core_shape(v, b)) for v in range(f.topo.nv)]
    dn.ctx.check(_capi.lib().qtt_eval_grad_host(dn.ctx.handle, dn.handle, _capi._dptr(pts), _capi._dptr(yb),
                                                pts.shape[0], _capi._dpp(grads), int(precision)))
    return grads


def apply_grad(op: TreeOperator, f: ITensorNetworkFunction, gbar, want_op: bool = True):
    """Pullback of operate(op, f) without truncation (the exact product W f):
    given the cotangent `gbar` of the product's cores (a list of arrays in the
    product's core shapes, e.g. eval_grad of the product), returns (fbar,
    wbar): the cotangents of f's cores and of the operator's cores (None when
    want_op is False)."""
    import torch
    dn, dop = device_network(f), device_operator(op)
    ctx = dn.ctx
    g = torch.tensor(np.concatenate([np.ravel(np.asarray(c, dtype=np.float64)) for c in gbar]), device="cuda")
    fb = torch.empty(dn.size(), dtype=torch.float64, device="cuda")
    wb = torch.empty(int(_capi.lib().qtt_mpo_size(dop.handle)), dtype=torch.float64, device="cuda") if want_op else None
    if g.numel() != sum(int(np.prod(s)) for s in _product_shapes(op, f)):
        raise ValueError("gbar must hold the exact product's cores")
    with _torch_stream(ctx, g):
        ctx.check(_capi.lib().qtt_apply_grad(ctx.handle, dop.handle, dn.handle, C.c_void_p(g.data_ptr()),
                                             C.c_void_p(fb.data_ptr()),
                                             C.c_void_p(wb.data_ptr()) if wb is not None else None))
    b = dn.bonds()
    fbar = _split(fb.cpu().numpy(), [f.topo.core_shape(v, b) for v in range(f.topo.nv)])
    wbar = None
    if want_op:
        ob = op.bonds()
        wbar = _split(wb.cpu().numpy(), [op.topo.op_shape(v, ob) for v in range(op.topo.nv)])
    return fbar, wbar


def _product_shapes(op, f):
    b = op.bonds() * device_network(f).bonds()
    return [f.topo.core_shape(v, b) for v in range(f.topo.nv)]


def _split(flat, shapes):
    out, o = [], 0
    for s in shapes:
        n = int(np.prod(s))
        out.append(flat[o:o + n].reshape(s))
        o += n
    return out


def operate_batch(ops, fs, cutoff=None, maxdim=None, precision: int = Precision.F64):
    """operate(ops[i], fs[i]; cutoff, maxdim) for a batch of independent functions:
    the truncation sweeps of the whole batch run in one launch."""
    n = len(fs)
    if len(ops) != n:
        raise ValueError("one operator per function")
    dns = [device_network(f) for f in fs]
    dops = [device_operator(o) for o in ops]
    ctx = engine_context()
    arr_op = (C.c_void_p * n)(*[d.handle.value for d in dops])
    arr_f = (C.c_void_p * n)(*[d.handle.value for d in dns])
    outs = (C.c_void_p * n)()
    errs = np.zeros(n)
    ctx.check(_capi.lib().qtt_apply_batch(ctx.handle, n, arr_op, arr_f, _cut(cutoff), int(maxdim or 0),
                                          int(precision), outs, _capi._dptr(errs)))
    res = [ITensorNetworkFunction(f.s, None, f.topo, device=_capi.DeviceNetwork(ctx, f.topo, C.c_void_p(h)))
           for f, h in zip(fs, outs)]
    return res, errs


def operate_batch_pipeline(ops, fs, xs_list=None, cutoff=None, maxdim=None, precision: int = Precision.F64,
                           take_sum: bool = False, truncate: bool = True):
    """One step of a batch of kernel-operator solves (the loop body of
    examples/fredholm_solver.jl:52-63 for many functions at once):
    g_i = operate(ops[i], fs[i]; cutoff, maxdim), its integral, and g_i at
    xs_list[i] (contiguous float64 CUDA tensors [npts_i, ndims], or None).
    The stages run back to back on the device, captured into one HIP graph,
    with one read-back at the end.  truncate=False: the exact products.
    Returns (gs, errs, integrals, values); an invalid point raises
    DigitMapError and no networks are returned."""
    import torch
    n = len(fs)
    if len(ops) != n:
        raise ValueError("one operator per function")
    dns = [device_network(f) for f in fs]
    dops = [device_operator(o) for o in ops]
    ctx = engine_context()
    arr_op = (C.c_void_p * n)(*[d.handle.value for d in dops])
    arr_f = (C.c_void_p * n)(*[d.handle.value for d in dns])
    outs = (C.c_void_p * n)()
    errs = np.zeros(n)
    ints = np.zeros(n)
    base = fs[0].s.base if n else 2
    vals = None
    arr_x = arr_v = npts_p = None
    if xs_list is not None:
        if len(xs_list) != n:
            raise ValueError("one point array per function")
        for f, x in zip(fs, xs_list):
            if not (isinstance(x, torch.Tensor) and x.is_cuda and x.dtype == torch.float64 and x.dim() == 2
                    and x.is_contiguous() and x.shape[1] == f.dimension()):
                raise ValueError("each point array must be a contiguous float64 CUDA tensor [npts, ndims]")
        vals = [torch.empty(x.shape[0], dtype=torch.float64, device=x.device) for x in xs_list]
        arr_x = (C.c_void_p * n)(*[x.data_ptr() for x in xs_list])
        arr_v = (C.c_void_p * n)(*[v.data_ptr() for v in vals])
        npts = np.array([x.shape[0] for x in xs_list], dtype=np.int64)
        npts_p = npts.ctypes.data_as(C.POINTER(C.c_int64))
    if n == 0:
        return [], errs, ints, vals
    lib = _capi.lib()
    cut = _cut(cutoff) if truncate else -1.0

    def call():
        return lib.qtt_batch_run(ctx.handle, n, arr_op, arr_f, cut, int(maxdim or 0), int(precision),
                                 1.0 if take_sum else 1.0 / base, arr_x, npts_p, arr_v, outs,
                                 _capi._dptr(errs), _capi._dptr(ints))
    if xs_list is not None:
        with _torch_stream(ctx, xs_list[0]):
            rc = call()
    else:
        rc = call()
    if rc == 2:  # QTT_ERANGE
        raise DigitMapError("a coordinate is negative or NaN (the digit map is defined for x >= 0)")
    ctx.check(rc)
    res = [ITensorNetworkFunction(f.s, None, f.topo, device=_capi.DeviceNetwork(ctx, f.topo, C.c_void_p(h)))
           for f, h in zip(fs, outs)]
    return res, errs, ints, vals


def evaluate_batch(fs, xs_list, precision: int = Precision.F64, outs=None):
    """evaluate(fs[i], xs_list[i]) for a batch of functions; each xs_list[i] is a
    contiguous float64 CUDA tensor [npts_i, ndims_i].  Returns the list of CUDA
    result tensors, queued on torch's current stream (as `evaluate`)."""
    import torch
    n = len(fs)
    if len(xs_list) != n:
        raise ValueError("one point array per function")
    dns = [device_network(f) for f in fs]
    for f, x in zip(fs, xs_list):
        if not (isinstance(x, torch.Tensor) and x.is_cuda and x.dtype == torch.float64 and x.dim() == 2
                and x.is_contiguous() and x.shape[1] == f.dimension()):
            raise ValueError("each point array must be a contiguous float64 CUDA tensor [npts, ndims]")
    if outs is None:
        outs = [torch.empty(x.shape[0], dtype=torch.float64, device=x.device) for x in xs_list]
    ctx = engine_context()
    arr_f = (C.c_void_p * n)(*[d.handle.value for d in dns])
    arr_x = (C.c_void_p * n)(*[x.data_ptr() for x in xs_list])
    arr_o = (C.c_void_p * n)(*[o.data_ptr() for o in outs])
    npts = np.array([x.shape[0] for x in xs_list], dtype=np.int64)
    if n == 0:
        return outs
    with _torch_stream(ctx, xs_list[0]):
        ctx.check(_capi.lib().qtt_eval_batch(ctx.handle, n, arr_f, arr_x,
                                             npts.ctypes.data_as(C.POINTER(C.c_int64)), arr_o, int(precision)))
    return outs


def integrate_batch(fs, take_sum: bool = False):
    """Integrals of a batch of functions in one launch."""
    n = len(fs)
    dns = [device_network(f) for f in fs]
    ctx = engine_context()
    arr = (C.c_void_p * n)(*[d.handle.value for d in dns])
    out = np.zeros(n)
    base = fs[0].s.base if n else 2
    ctx.check(_capi.lib().qtt_integrate_batch(ctx.handle, n, arr, 1.0 if take_sum else 1.0 / base,
                                              _capi._dptr(out)))
    return out
