"""CPU: host-side bookkeeping of minirender_amd.Server (no GPU).

A wait that times out leaves the request live in the C server
(mr_server_wait keeps the ticket on MR_ERR_TIMEOUT), whose worker will still
write the request's outputs: the Python side must keep those arrays alive
until a wait that completes.  The C library is replaced by a stub that
answers MR_ERR_TIMEOUT once, then MR_OK.
"""
import gc

import numpy as np
import pytest

import minirender_amd as mra
from minirender_amd import MrError


class _StubLib:
    def __init__(self, codes):
        self.codes = list(codes)
        self.calls = []

    def mr_server_wait(self, h, ticket, timeout_ms, info):
        self.calls.append((ticket, timeout_ms))
        return self.codes.pop(0)

    def mr_server_destroy(self, h):
        return 0


def _server(codes):
    srv = mra.Server.__new__(mra.Server)
    srv.lib = _StubLib(codes)
    srv.h = None  # close() is then a no-op
    srv._pending = {}
    return srv


def test_timeout_keeps_outputs_until_completed_wait():
    srv = _server([-9, 0])
    out = {"rgb8": np.zeros((4, 4, 3), np.uint8)}
    srv._pending[7] = (out, None, None)
    del out
    gc.collect()
    with pytest.raises(MrError) as e:
        srv.wait(7, timeout_ms=1)
    assert e.value.code == -9
    assert 7 in srv._pending, "timed-out request's outputs were released"
    got, info = srv.wait(7)
    assert got["rgb8"].shape == (4, 4, 3)
    assert 7 not in srv._pending
    assert srv.lib.calls == [(7, 1), (7, -1)]


def test_failed_request_is_forgotten():
    # Any status other than a timeout means the C server forgot the ticket.
    srv = _server([-8])
    srv._pending[3] = ({}, None, None)
    with pytest.raises(MrError):
        srv.wait(3)
    assert 3 not in srv._pending
