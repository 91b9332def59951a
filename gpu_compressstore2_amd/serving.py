"""Python view of the continuous-batching server and the HBM-budget arena of
libcsx.so (csx_server_*, csx_arena_*).  The scheduling itself is C++."""
import ctypes

from . import CsxError, _check, lib

_vp = ctypes.c_void_p
_u64 = ctypes.c_uint64
_u32 = ctypes.c_uint32


class ServerConfig(ctypes.Structure):
    _fields_ = [("streams", _u32), ("max_batch", _u32), ("max_batch_elements", _u64)]


class Request(ctypes.Structure):
    _fields_ = [("id", _u64), ("input", _vp), ("mask", _vp), ("out", _vp), ("n", _u64), ("elem_size", _u32), ("ready", _vp)]


class Completion(ctypes.Structure):
    _fields_ = [("id", _u64), ("count", _u64), ("latency_ms", ctypes.c_double)]


def _bind():
    return lib()


class Arena:
    """Device memory budget carved first-fit (csx_arena_*)."""

    def __init__(self, nbytes):
        self.L = _bind()
        self.h = _vp()
        _check(self.L.csx_arena_create(nbytes, ctypes.byref(self.h)))

    def alloc(self, nbytes):
        """Device address, or None when the budget has no room."""
        p = _vp()
        st = self.L.csx_arena_alloc(self.h, nbytes, ctypes.byref(p))
        if st == 2:
            return None
        _check(st)
        return p.value

    def free(self, ptr):
        _check(self.L.csx_arena_free(self.h, _vp(ptr)))

    def used(self):
        return int(self.L.csx_arena_used(self.h))

    def close(self):
        if self.h:
            _check(self.L.csx_arena_destroy(self.h))
            self.h = None


class Server:
    """Continuous batching over device-resident requests."""

    def __init__(self, streams=4, max_batch=256, max_batch_elements=1 << 32):
        self.L = _bind()
        self.h = _vp()
        cfg = ServerConfig(streams, max_batch, max_batch_elements)
        _check(self.L.csx_server_create(_vp(ctypes.addressof(cfg)), ctypes.byref(self.h)))
        self._done = (Completion * 4096)()

    def submit(self, rid, inp, mask, out, n=None, ready=None):
        n = inp.numel() if n is None else n
        r = Request(rid, inp.data_ptr(), mask.data_ptr(), out.data_ptr(), n, inp.element_size(), ready)
        _check(self.L.csx_server_submit(self.h, _vp(ctypes.addressof(r))))

    def step(self):
        k = _u32(0)
        _check(self.L.csx_server_step(self.h, ctypes.byref(k)))
        return k.value

    def poll(self):
        k = _u32(0)
        _check(self.L.csx_server_poll(self.h, ctypes.cast(self._done, _vp), len(self._done), ctypes.byref(k)))
        return [(self._done[i].id, self._done[i].count, self._done[i].latency_ms) for i in range(k.value)]

    def queued(self):
        return int(self.L.csx_server_queued(self.h))

    def inflight(self):
        return int(self.L.csx_server_inflight(self.h))

    def drain(self):
        out = []
        while self.queued() or self.inflight():
            self.step()
            out += self.poll()
        return out

    def close(self):
        if self.h:
            _check(self.L.csx_server_destroy(self.h))
            self.h = None


__all__ = ["Arena", "Server", "ServerConfig", "Request", "Completion"]
