"""Serving scheduler on the host (simulated device): admission by the memory
budget, FIFO order, waiting for memory, deadlines, cancellation, scenes that
are still building, and statistics.  No GPU needed: failures are forced from
the host (a small injected budget, a mock of the device's service time)."""
import time

import pytest

from igx_raytracing_amd import RtError, Server
from igx_raytracing_amd.serve import DONE, FAILED

RT_ERR_ADMISSION = -5
RT_ERR_CANCELLED = -6
RAY_BYTES = 32


def closest_bytes(n):
    up = lambda v: (v + 255) // 256 * 256  # noqa: E731
    return up(n * RAY_BYTES) + up(n * 16)


def test_refuses_request_larger_than_budget():
    s = Server(simulate=True, mem_budget=closest_bytes(1000))
    sid = s.add_scene()
    with pytest.raises(RtError) as e:
        s.submit(sid, "closest", 1001)
    assert e.value.code == RT_ERR_ADMISSION
    t = s.submit(sid, "closest", 1000)
    assert s.wait(t, 2000)[0] == DONE
    s.close()


def test_waits_for_memory_and_keeps_fifo_order():
    # Budget for exactly one 100k-ray request: the second waits for the first.
    s = Server(simulate=True, mem_budget=closest_bytes(100_000), streams=4)
    sid = s.add_scene()
    tickets = [s.submit(sid, "closest", 100_000) for _ in range(5)]
    done_at = []
    for t in tickets:
        st, status, lat = s.wait(t, 5000)
        assert st == DONE and status == 0
        done_at.append(lat)
    assert done_at == sorted(done_at), "FIFO completion under a one-request budget"
    stats = s.stats()
    assert stats["ticks"] == 5 and stats["mean_requests_per_tick"] == 1.0
    assert stats["mem_peak"] <= stats["mem_budget"]
    s.close()


def test_batches_requests_that_arrive_together():
    s = Server(simulate=True, mem_budget=1 << 30, streams=1)
    sid = s.add_scene()
    first = s.submit(sid, "closest", 2_000_000)  # ~2 ms modelled
    rest = [s.submit(sid, "occluded", 1000) for _ in range(20)]
    for t in [first] + rest:
        assert s.wait(t, 5000)[0] == DONE
    assert s.stats()["ticks"] < 21, "requests waiting behind a tick share the next one"
    s.close()


def test_plain_mode_serves_one_request_per_launch():
    s = Server(simulate=True, mem_budget=1 << 30, plain=True)
    sid = s.add_scene()
    ts = [s.submit(sid, "closest", 1000) for _ in range(10)]
    for t in ts:
        assert s.wait(t, 5000)[0] == DONE
    st = s.stats()
    assert st["ticks"] == 10 and st["mean_requests_per_tick"] == 1.0
    s.close()


def test_deadline_and_cancel():
    s = Server(simulate=True, mem_budget=closest_bytes(3_000_000), streams=1)
    sid = s.add_scene()
    big = s.submit(sid, "closest", 3_000_000)          # holds the whole budget ~3 ms
    late = s.submit(sid, "closest", 10, deadline_ms=0.5)  # cannot start in time
    gone = s.submit(sid, "closest", 10)
    s.cancel(gone)
    assert s.poll(gone)[:2] == (FAILED, RT_ERR_CANCELLED)
    assert s.wait(big, 5000)[0] == DONE
    st, status, _ = s.wait(late, 5000)
    assert (st, status) == (FAILED, RT_ERR_CANCELLED)
    with pytest.raises(RtError):
        s.cancel(big)  # already finished
    s.forget(big)
    with pytest.raises(RtError):
        s.poll(big)
    st = s.stats()
    assert st["failed"] == 2 and st["completed"] == 1
    s.close()


def test_unknown_scene_and_zero_rays():
    s = Server(simulate=True)
    with pytest.raises(RtError):
        s.submit(0, "closest", 10)
    sid = s.add_scene()
    t = s.submit(sid, "closest", 0)
    assert s.poll(t)[0] == DONE
    s.close()


def test_server_stops_cleanly_with_work_in_flight():
    s = Server(simulate=True, mem_budget=1 << 30, streams=2)
    sid = s.add_scene()
    for _ in range(50):
        s.submit(sid, "occluded", 50_000)
    t0 = time.time()
    s.close()  # drains pending and in-flight ticks
    assert time.time() - t0 < 10


def test_failed_tick_completions_fail_their_requests_and_do_not_hang():
    # Every tick completes with a HIP error (a device that faulted): each
    # request ends FAILED with RT_ERR_HIP, waits return, and the server shuts
    # down, instead of polling the broken ticks forever.
    RT_ERR_HIP = -3
    s = Server(simulate="fail", mem_budget=closest_bytes(400_000), streams=2)
    sid = s.add_scene()
    tickets = [s.submit(sid, "closest" if k % 2 else "occluded", 50_000 + k) for k in range(12)]
    t0 = time.perf_counter()
    for t in tickets:
        st, status, _ = s.wait(t, 5000)
        assert st == FAILED and status == RT_ERR_HIP
    assert time.perf_counter() - t0 < 5.0
    stats = s.stats()
    assert stats["failed"] == 12 and stats["completed"] == 0
    # The memory of failed requests was released: a request needing the whole
    # budget is still admitted.
    t = s.submit(sid, "closest", 400_000)
    assert s.wait(t, 5000)[0] == FAILED
    s.close()
