"""Stream/event pool names and argument checks that need no GPU (the pools themselves are
tested in test_gpu_streams.py)."""
import pytest

import curadon_amd as cm


def test_reference_names_exported():
    # python/standalone/src/curadon/__init__.py:12
    for name in ("event", "event_view", "event_pool", "stream", "stream_view", "stream_pool",
                 "get_next_stream", "get_stream", "get_next_event", "get_event"):
        assert hasattr(cm, name), name


@pytest.mark.parametrize("cls", ["stream_pool", "event_pool"])
def test_empty_pool_rejected(cls):
    with pytest.raises(ValueError):
        getattr(cm, cls)(0)
