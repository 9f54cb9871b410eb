"""GPUAsyncProcess: streams, finish, generic batched_run (cuvarbase/core.py:13-56)."""
from __future__ import annotations

from ._device import Stream, set_device, synchronize


class GPUAsyncProcess(object):
    """Base of the asynchronous GPU processes.

    Keyword arguments as cuvarbase/core.py:14-25: ``reader``, ``nstreams``,
    ``function_kwargs``, ``device``.  The device context is created lazily
    through the C ABI (no import-time context).
    """

    def __init__(self, *args, **kwargs):
        self.reader = kwargs.get('reader', None)
        self.nstreams = kwargs.get('nstreams', None)
        self.function_kwargs = kwargs.get('function_kwargs', {})
        self.device = kwargs.get('device', 0)
        self.streams = []
        self.gpu_data = []
        self.results = []
        self._adjust_nstreams = self.nstreams is None
        self._inflight = []
        if self.nstreams is not None:
            self._create_streams(self.nstreams)
        self.prepared_functions = {}

    def _select_device(self):
        set_device(self.device)

    def _create_streams(self, n):
        self._select_device()
        for _ in range(n):
            self.streams.append(Stream())

    def _compile_and_prepare_functions(self):
        raise NotImplementedError()

    def run(self, *args, **kwargs):
        raise NotImplementedError()

    def finish(self):
        """Synchronise all active streams (cuvarbase/core.py:37-40)."""
        for stream in self.streams:
            stream.synchronize()
        # memory objects created with stream=None use the null stream
        synchronize()
        # buffers of finished asynchronous batches can be released now
        self._inflight = []

    def batched_run(self, data, batch_size=10, **kwargs):
        """Run ``data`` in batches of ``batch_size`` (cuvarbase/core.py:42-56)."""
        nsubmit = 0
        results = []
        while nsubmit < len(data):
            batch = []
            while len(batch) < batch_size and nsubmit < len(data):
                batch.append(data[nsubmit])
                nsubmit += 1
            res = self.run(batch, **kwargs)
            self.finish()
            results.extend(res)
        return results
