"""Host-side failure handling of the serving loop (no GPU needed): a server
none of whose slots can start completes every request with an error instead
of leaving its waiter blocked, and malformed requests are rejected at submit,
before they could share a batch.  The slots here ask for a device index that
does not exist, so the same failure is forced on any machine."""
import numpy as np
import pytest

from strsets import pack

NO_SUCH_DEVICE = 4096


@pytest.mark.timeout(120)
def test_server_without_slots_fails_requests(ps5):
    srv = ps5.Server(NO_SUCH_DEVICE, hbm_budget=1 << 30, slots=3)
    try:
        arena, offs = pack([b"b", b"a", b"c"])
        ids = [srv.submit(arena, offs) for _ in range(4)]
        for rid in ids:
            with pytest.raises(ps5.Ps5Error) as e:
                srv.wait(rid)
            assert "no serving slot could start" in str(e.value)
        h = srv.health()
        assert h["live_slots"] == 0 and h["failed"] == 4
        # a request submitted after the slots failed fails at once as well
        rid = srv.submit(arena, offs)
        with pytest.raises(ps5.Ps5Error):
            srv.wait(rid)
    finally:
        srv.close()


@pytest.mark.timeout(120)
def test_server_rejects_malformed_requests(ps5):
    srv = ps5.Server(NO_SUCH_DEVICE, hbm_budget=1 << 30, slots=1)
    try:
        arena, offs = pack([b"xy", b"z"])
        bad = offs.copy()
        bad[1] = arena.size          # an offset outside the arena
        with pytest.raises(ps5.Ps5Error, match="outside the arena"):
            srv.submit(arena, bad)
        unterminated = np.frombuffer(b"xy\0zz", np.uint8).copy()
        with pytest.raises(ps5.Ps5Error, match="last byte must be a NUL"):
            srv.submit(unterminated, np.array([0, 3], np.uint32))
        # rejected requests take no id: the first valid one gets id 1, after
        # which a fault can no longer be planted in it
        srv.inject_fault(2)
        rid = srv.submit(arena, offs)
        assert rid == 1
        with pytest.raises(ps5.Ps5Error, match="already submitted"):
            srv.inject_fault(1)
    finally:
        srv.close()
