"""GPUAsyncProcess: the device, streams and batch loop shared by the GLS and
NFFT processes.

Public surface of cuvarbase/core.py:13-56: the keyword arguments ``reader``,
``nstreams``, ``function_kwargs`` and ``device``; the attributes ``streams``
and ``prepared_functions``; ``run``, ``finish`` and ``batched_run``.  The
device context is created lazily through the C ABI (the reference creates
one when the package is imported).
"""
from __future__ import annotations

from ._device import Stream, set_device, synchronize

# keyword arguments of every process and their defaults
PROCESS_KEYWORDS = (("reader", None), ("nstreams", None), ("function_kwargs", None),
                    ("device", 0))


class GPUAsyncProcess(object):
    """Base class of the asynchronous processes.

    Work is enqueued on ``streams`` and is complete after ``finish()``.
    Device buffers of batches still in flight are kept alive until then, and
    results that need host post-processing (an upcast, a copy out of staging
    memory) are finished by callbacks that ``finish()`` runs after the
    streams have drained.
    """

    def __init__(self, *args, **kwargs):
        for name, default in PROCESS_KEYWORDS:
            setattr(self, name, kwargs.get(name, default))
        if self.function_kwargs is None:
            self.function_kwargs = {}
        self.streams = []
        self.prepared_functions = {}
        self._keep_alive = []       # device objects used by enqueued work
        self._on_finish = []        # host steps that need the enqueued work done
        if self.nstreams:
            self._create_streams(self.nstreams)

    def _select_device(self):
        set_device(self.device)

    def _create_streams(self, count):
        self._select_device()
        self.streams.extend(Stream() for _ in range(int(count)))

    def _stream(self, i=0):
        """Stream i of the process, creating streams up to it."""
        if len(self.streams) <= i:
            self._create_streams(i + 1 - len(self.streams))
        return self.streams[i]

    def _compile_and_prepare_functions(self, **kwargs):
        raise NotImplementedError()

    def run(self, *args, **kwargs):
        raise NotImplementedError()

    def finish(self):
        """Wait for all enqueued work (every stream of the process and the
        null stream, which memory objects made without a stream use), then
        complete the host side of the results."""
        for stream in self.streams:
            stream.synchronize()
        synchronize()
        steps, self._on_finish = self._on_finish, []
        for step in steps:
            step()
        self._keep_alive = []

    def batched_run(self, data, batch_size=10, **kwargs):
        """``run`` over consecutive slices of at most ``batch_size`` items,
        each finished before the next is enqueued; the results of all slices
        in order."""
        if batch_size < 1:
            raise ValueError("batch_size must be at least 1")
        results = []
        for lo in range(0, len(data), batch_size):
            results.extend(self.run(data[lo:lo + batch_size], **kwargs))
            self.finish()
        return results
