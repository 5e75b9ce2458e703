"""Stream and event pools with the names curadon exports (python/standalone/src/curadon/__init__.py:12,
bound in python/standalone/src/stream.cu:7-36 over include/curadon/pool.hpp:16-228).

Every projector entry point of this package takes the caller's stream (torch's current stream
from Python, a ``void *stream`` at the C-ABI), so the pools are conveniences for code that
imports them: a ``stream`` owns a HIP stream (a ``torch.cuda.Stream`` at default priority), an
``event`` owns a HIP event, the ``*_view`` classes wrap one without owning it, and the pools hand
them out round-robin. ``value()`` returns the raw handle (an int) for the C-ABI; ``torch``
returns the torch object for ``with torch.cuda.stream(...)``.

Semantics follow pool.hpp: ``stream_pool(n)`` / ``event_pool(n)`` create n objects,
``get_next_*`` cycles through them (an atomic counter there, a lock here), ``get_*(idx)`` takes
``idx % n``; the module-level ``get_next_stream`` / ``get_stream`` / ``get_next_event`` /
``get_event`` use process-wide pools of 16, created on first use.
"""
import itertools
import threading

import torch

__all__ = ["event", "event_view", "event_pool", "stream", "stream_view", "stream_pool",
           "get_next_stream", "get_stream", "get_next_event", "get_event"]

DEFAULT_POOL_SIZE = 16  # pool.hpp:137,174


class stream_view:
    """Non-owning view of a HIP stream (pool.hpp:92-121)."""

    def __init__(self, s=None):
        # None views the current stream of the current device, as a default-constructed
        # stream_view views the default stream
        self._s = torch.cuda.current_stream() if s is None else s

    @property
    def torch(self):
        return self._s

    def value(self):
        return self._s.cuda_stream

    def synchronize(self):
        self._s.synchronize()

    def wait_for_event(self, ev):
        self._s.wait_event(_torch_event(ev))


class stream(stream_view):
    """Owning HIP stream (pool.hpp:64-90): created at default priority."""

    def __init__(self):
        super().__init__(torch.cuda.Stream())


class event_view:
    """Non-owning view of a HIP event (pool.hpp:40-62)."""

    def __init__(self, e):
        self._e = e

    @property
    def torch(self):
        return self._e

    def value(self):
        return self._e.cuda_event

    def record(self, s=None):
        self._e.record(_torch_stream(s))

    def synchronize(self):
        self._e.synchronize()


class event(event_view):
    """Owning HIP event (pool.hpp:16-38)."""

    def __init__(self):
        e = torch.cuda.Event()
        e.record()  # torch creates the HIP event on first record; value() needs the handle
        super().__init__(e)


def _torch_stream(s):
    if s is None:
        return torch.cuda.current_stream()
    return s.torch if isinstance(s, stream_view) else s


def _torch_event(e):
    return e.torch if isinstance(e, event_view) else e


class _Pool:
    def __init__(self, make, view, pool_size):
        if int(pool_size) < 1:
            raise ValueError("pool size must be at least 1")
        self._items = [make() for _ in range(int(pool_size))]
        self._view = view
        self._next = itertools.count()
        self._lock = threading.Lock()

    def _get_next(self):
        with self._lock:
            i = next(self._next)
        return self._view(self._items[i % len(self._items)].torch)

    def _get(self, idx):
        return self._view(self._items[int(idx) % len(self._items)].torch)

    def __len__(self):
        return len(self._items)


class stream_pool(_Pool):
    """Round-robin pool of owned streams (pool.hpp:135-168)."""

    def __init__(self, pool_size=DEFAULT_POOL_SIZE):
        super().__init__(stream, stream_view, pool_size)

    def get_next_stream(self):
        return self._get_next()

    def get_stream(self, idx):
        return self._get(idx)


class event_pool(_Pool):
    """Round-robin pool of owned events (pool.hpp:170-199)."""

    def __init__(self, pool_size=DEFAULT_POOL_SIZE):
        super().__init__(event, event_view, pool_size)

    def get_next_event(self):
        return self._get_next()

    def get_event(self, idx):
        return self._get(idx)


_pools = {}
_pools_lock = threading.Lock()


def _instance(kind):
    with _pools_lock:
        p = _pools.get(kind)
        if p is None:
            p = stream_pool() if kind == "stream" else event_pool()
            _pools[kind] = p
        return p


def get_next_stream():
    return _instance("stream").get_next_stream()


def get_stream(idx):
    return _instance("stream").get_stream(idx)


def get_next_event():
    return _instance("event").get_next_event()


def get_event(idx):
    return _instance("event").get_event(idx)
