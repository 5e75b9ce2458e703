"""Stream and event pools under curadon's names (python/standalone/src/stream.cu:7-36,
include/curadon/pool.hpp:135-228): round-robin order, views of owned objects, event ordering
between pool streams, and projections issued on a pool stream."""
import numpy as np
import pytest
import torch

import curadon_amd as cm
from geoms import as_fan_geometry, fan

pytestmark = pytest.mark.gpu


def test_pools_round_robin(gpu):
    sp = cm.stream_pool(3)
    got = [sp.get_next_stream().value() for _ in range(7)]
    assert got[:3] == got[3:6] and got[6] == got[0] and len(set(got[:3])) == 3
    assert sp.get_stream(4).value() == got[1]
    ep = cm.event_pool(2)
    ev = [ep.get_next_event().value() for _ in range(4)]
    assert ev[0] == ev[2] and ev[1] == ev[3] and ev[0] != ev[1]
    assert cm.get_stream(5).value() == cm.get_stream(21).value()  # pool of 16
    assert cm.get_event(0).value() == cm.get_event(16).value()
    assert all(s.priority == 0 for s in (sp.get_stream(i).torch for i in range(3)))


def test_projection_on_pool_streams_with_event_ordering(gpu):
    geom = as_fan_geometry(fan(64, 91, 90, 128.0, 192.0))
    x = torch.rand(64, 64, device=gpu)
    ref = cm.forward(x, geom)
    torch.cuda.synchronize()
    s1, s2 = cm.get_next_stream(), cm.get_next_stream()
    ev = cm.get_next_event()
    with torch.cuda.stream(s1.torch):
        y = cm.forward(x, geom)
        ev.record(s1)
    s2.wait_for_event(ev)
    with torch.cuda.stream(s2.torch):
        z = cm.backward(y, geom)
    s2.synchronize()
    assert torch.equal(y, ref)
    assert torch.equal(z, cm.backward(ref, geom))
