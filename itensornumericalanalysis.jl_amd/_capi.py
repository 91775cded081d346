"""ctypes binding of libqtt.so (include/qtt.h).  The engine is the only compute
path: if the library is missing or no GPU is visible these calls raise."""
from __future__ import annotations

import atexit
import ctypes as C
import os
import threading
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libqtt.so")

QTT_OK, QTT_EINVAL, QTT_ERANGE, QTT_ENOMEM, QTT_EHIP, QTT_ENOCONV = 0, 1, 2, 3, 4, 5
_NAMES = {1: "EINVAL", 2: "ERANGE", 3: "ENOMEM", 4: "EHIP", 5: "ENOCONV"}

# every symbol declared in include/qtt.h
EXPORTS = [
    "qtt_version", "qtt_ctx_create", "qtt_ctx_destroy", "qtt_last_error", "qtt_ctx_set_stream",
    "qtt_sync", "qtt_net_create", "qtt_net_bonds", "qtt_net_read", "qtt_net_copy",
    "qtt_net_destroy", "qtt_mpo_create", "qtt_mpo_destroy", "qtt_eval", "qtt_eval_host",
    "qtt_integrate", "qtt_inner", "qtt_apply", "qtt_truncate", "qtt_eval_grad",
    "qtt_eval_grad_host", "qtt_net_size", "qtt_apply_batch", "qtt_integrate_batch", "qtt_eval_batch",
    "qtt_tune_eval", "qtt_eval_noalloc", "qtt_apply_grad", "qtt_mpo_size", "qtt_batch_run",
    "qtt_multiply", "qtt_add", "qtt_partial_integrate", "qtt_operator_proj", "qtt_mpo_bonds", "qtt_mpo_read",
    "qtt_net_save", "qtt_net_load", "qtt_mpo_save", "qtt_mpo_load", "qtt_bp_eval_host",
]


class EngineError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libqtt {_NAMES.get(code, code)}: {msg}")
        self.code = code


class QttGraph(C.Structure):
    _fields_ = [("nv", C.c_int32), ("ne", C.c_int32), ("eu", C.POINTER(C.c_int32)), ("ev", C.POINTER(C.c_int32)),
                ("ebond", C.POINTER(C.c_int64)), ("sdim", C.POINTER(C.c_int32)), ("vcoord", C.POINTER(C.c_int32)),
                ("vdigit", C.POINTER(C.c_int32)), ("ncoord", C.c_int32), ("ndigits", C.POINTER(C.c_int32))]


class QttTopology(C.Structure):
    _fields_ = [("nv", C.c_int32), ("parent", C.POINTER(C.c_int32)), ("sdim", C.POINTER(C.c_int32)),
                ("vcoord", C.POINTER(C.c_int32)), ("vdigit", C.POINTER(C.c_int32)),
                ("ncoord", C.c_int32), ("ndigits", C.POINTER(C.c_int32))]


_lib = None
_lock = threading.Lock()


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise EngineError(QTT_EINVAL, f"{LIBPATH} is not built (run __graft_entry__.build())")
        L = C.CDLL(LIBPATH)
        vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int64)
        tp = C.POINTER(QttTopology)
        dpp = C.POINTER(dp)
        L.qtt_version.restype = C.c_int
        L.qtt_ctx_create.argtypes = [C.c_int, C.POINTER(vp)]
        L.qtt_ctx_destroy.argtypes = [vp]
        L.qtt_ctx_destroy.restype = None
        L.qtt_last_error.argtypes = [vp]
        L.qtt_last_error.restype = C.c_char_p
        L.qtt_ctx_set_stream.argtypes = [vp, vp]
        L.qtt_sync.argtypes = [vp]
        L.qtt_net_create.argtypes = [vp, tp, ip, dpp, C.POINTER(vp)]
        L.qtt_net_bonds.argtypes = [vp, vp, ip]
        L.qtt_net_read.argtypes = [vp, vp, dpp]
        L.qtt_net_copy.argtypes = [vp, vp, C.POINTER(vp)]
        L.qtt_net_destroy.argtypes = [vp]
        L.qtt_net_destroy.restype = None
        L.qtt_net_size.argtypes = [vp]
        L.qtt_net_size.restype = C.c_int64
        L.qtt_mpo_create.argtypes = [vp, tp, ip, dpp, C.POINTER(vp)]
        L.qtt_mpo_destroy.argtypes = [vp]
        L.qtt_mpo_destroy.restype = None
        L.qtt_eval.argtypes = [vp, vp, vp, C.c_int64, vp, C.c_int]
        L.qtt_eval_host.argtypes = [vp, vp, dp, C.c_int64, dp, C.c_int]
        L.qtt_integrate.argtypes = [vp, vp, C.c_double, dp]
        L.qtt_inner.argtypes = [vp, vp, vp, vp, dp]
        L.qtt_apply.argtypes = [vp, vp, vp, C.c_double, C.c_int64, C.c_int, C.POINTER(vp), dp]
        L.qtt_truncate.argtypes = [vp, vp, C.c_double, C.c_int64, dp]
        L.qtt_eval_grad.argtypes = [vp, vp, vp, vp, C.c_int64, vp, C.c_int]
        L.qtt_eval_grad_host.argtypes = [vp, vp, dp, dp, C.c_int64, dpp, C.c_int]
        L.qtt_apply_batch.argtypes = [vp, C.c_int, C.POINTER(vp), C.POINTER(vp), C.c_double, C.c_int64, C.c_int,
                                      C.POINTER(vp), dp]
        L.qtt_integrate_batch.argtypes = [vp, C.c_int, C.POINTER(vp), C.c_double, dp]
        L.qtt_eval_batch.argtypes = [vp, C.c_int, C.POINTER(vp), C.POINTER(vp), ip, C.POINTER(vp), C.c_int]
        i32p = C.POINTER(C.c_int32)
        L.qtt_multiply.argtypes = [vp, vp, vp, C.POINTER(QttTopology), i32p, i32p, C.POINTER(vp)]
        L.qtt_add.argtypes = [vp, vp, vp, C.POINTER(vp)]
        L.qtt_operator_proj.argtypes = [vp, vp, C.POINTER(vp)]
        L.qtt_mpo_bonds.argtypes = [vp, vp, ip]
        L.qtt_mpo_read.argtypes = [vp, vp, dpp]
        cpp = C.POINTER(C.c_char_p)
        L.qtt_net_save.argtypes = [vp, vp, C.c_char_p, cpp]
        L.qtt_net_load.argtypes = [vp, C.c_char_p, C.POINTER(vp)]
        L.qtt_mpo_save.argtypes = [vp, vp, C.c_char_p, cpp]
        L.qtt_mpo_load.argtypes = [vp, C.c_char_p, C.POINTER(vp)]
        L.qtt_bp_eval_host.argtypes = [vp, C.POINTER(QttGraph), dpp, dp, C.c_int64, dp, C.c_int, C.c_double,
                                       C.POINTER(C.c_int)]
        L.qtt_partial_integrate.argtypes = [vp, vp, i32p, C.c_double, C.POINTER(QttTopology), i32p,
                                            C.POINTER(vp)]
        L.qtt_batch_run.argtypes = [vp, C.c_int, C.POINTER(vp), C.POINTER(vp), C.c_double, C.c_int64, C.c_int,
                                    C.c_double, C.POINTER(vp), ip, C.POINTER(vp), C.POINTER(vp), dp, dp]
        L.qtt_tune_eval.argtypes = [vp, vp, vp, C.c_int64, vp, C.POINTER(C.c_int), dp]
        L.qtt_eval_noalloc.argtypes = [vp, vp, vp, C.c_int64, vp]
        L.qtt_apply_grad.argtypes = [vp, vp, vp, vp, vp, vp]
        L.qtt_mpo_size.argtypes = [vp]
        L.qtt_mpo_size.restype = C.c_int64
        _lib = L
    return _lib


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _dpp(arrs):
    return (C.POINTER(C.c_double) * len(arrs))(*[_dptr(a) for a in arrs])


class Context:
    """A libqtt context (one HIP stream) on one device."""

    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        rc = lib().qtt_ctx_create(int(device), C.byref(self._h))
        if rc:
            raise EngineError(rc, "qtt_ctx_create failed (no HIP device visible?)")
        self.device = device
        self.stream = None  # the HIP stream work runs on (None: the context's own)

    @property
    def handle(self):
        return self._h

    def check(self, rc):
        if rc:
            msg = lib().qtt_last_error(self._h)
            raise EngineError(rc, msg.decode() if msg else "")

    def set_stream(self, stream_ptr):
        self.check(lib().qtt_ctx_set_stream(self._h, C.c_void_p(stream_ptr) if stream_ptr else None))
        self.stream = stream_ptr or None

    def sync(self):
        self.check(lib().qtt_sync(self._h))

    def close(self):
        if self._h:
            lib().qtt_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        if _shutting_down:
            return
        try:
            self.close()
        except Exception:
            pass


# Device handles still alive.  At interpreter exit they are all destroyed,
# networks and operators first, then the context, while the HIP runtime is
# still up; the destructors that the garbage collector runs later do nothing.
_live = weakref.WeakSet()
_shutting_down = False


def release_all():
    """Destroy every live network and operator handle, then the process-wide
    context (registered with atexit; callable earlier by a program that is done
    with the engine)."""
    global _shutting_down, _ctx
    _shutting_down = True
    for obj in list(_live):
        try:
            obj._release()
        except Exception:
            pass
    ctx, _ctx = _ctx, None
    if ctx is not None:
        try:
            ctx.close()
        except Exception:
            pass


atexit.register(release_all)

_ctx = None


def context() -> Context:
    global _ctx
    with _lock:
        if _ctx is None:
            _ctx = Context(0)
        return _ctx


def _topology_struct(topo):
    keep = [np.ascontiguousarray(a, dtype=np.int32) for a in
            (topo.parent, topo.sdim, topo.vcoord, topo.vdigit, topo.Lc if len(topo.Lc) else np.zeros(1))]
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))
    st = QttTopology(int(topo.nv), ip(keep[0]), ip(keep[1]), ip(keep[2]), ip(keep[3]), int(topo.ncoord),
                     ip(keep[4]))
    return st, keep


class DeviceNetwork:
    """Cores of a tree network resident on the device (qtt_net)."""

    def __init__(self, ctx: Context, topo, handle):
        self.ctx = ctx
        self.topo = topo
        self._h = handle
        _live.add(self)

    @classmethod
    def upload(cls, ctx: Context, topo, cores):
        st, keep = _topology_struct(topo)
        cores = [np.ascontiguousarray(c, dtype=np.float64) for c in cores]
        bond = np.array([c.shape[0] for c in cores], dtype=np.int64)
        h = C.c_void_p()
        ctx.check(lib().qtt_net_create(ctx.handle, C.byref(st), bond.ctypes.data_as(C.POINTER(C.c_int64)),
                                       _dpp(cores), C.byref(h)))
        return cls(ctx, topo, h)

    @property
    def handle(self):
        return self._h

    def bonds(self):
        b = np.zeros(self.topo.nv, dtype=np.int64)
        self.ctx.check(lib().qtt_net_bonds(self.ctx.handle, self._h, b.ctypes.data_as(C.POINTER(C.c_int64))))
        return b

    def size(self):
        return int(lib().qtt_net_size(self._h))

    def download(self):
        b = self.bonds()
        cores = [np.empty(self.topo.core_shape(v, b)) for v in range(self.topo.nv)]
        self.ctx.check(lib().qtt_net_read(self.ctx.handle, self._h, _dpp(cores)))
        return cores

    def copy(self):
        h = C.c_void_p()
        self.ctx.check(lib().qtt_net_copy(self.ctx.handle, self._h, C.byref(h)))
        return DeviceNetwork(self.ctx, self.topo, h)

    def _release(self):
        if self._h:
            lib().qtt_net_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        if _shutting_down:
            return
        try:
            self._release()
        except Exception:
            pass


class DeviceOperator:
    """An operator network resident on the device (qtt_mpo)."""

    def __init__(self, ctx: Context, topo, cores=None, handle=None):
        self.ctx = ctx
        self.topo = topo
        self._h = C.c_void_p()
        if handle is not None:
            self._h = handle
        else:
            st, keep = _topology_struct(topo)
            cores = [np.ascontiguousarray(c, dtype=np.float64) for c in cores]
            bond = np.array([c.shape[0] for c in cores], dtype=np.int64)
            ctx.check(lib().qtt_mpo_create(ctx.handle, C.byref(st), bond.ctypes.data_as(C.POINTER(C.c_int64)),
                                           _dpp(cores), C.byref(self._h)))
        _live.add(self)

    def bonds(self):
        b = np.zeros(self.topo.nv, dtype=np.int64)
        self.ctx.check(lib().qtt_mpo_bonds(self.ctx.handle, self._h, b.ctypes.data_as(C.POINTER(C.c_int64))))
        return b

    def download(self):
        b = self.bonds()
        cores = [np.empty(self.topo.op_shape(v, b)) for v in range(self.topo.nv)]
        self.ctx.check(lib().qtt_mpo_read(self.ctx.handle, self._h, _dpp(cores)))
        return cores

    @property
    def handle(self):
        return self._h

    def _release(self):
        if self._h:
            lib().qtt_mpo_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        if _shutting_down:
            return
        try:
            self._release()
        except Exception:
            pass
