"""The serving loop: continuous batching over a request queue with
admission against a memory budget.  Each request's probe series must be the
same alone and in a mix, bit for bit, and equal to a single context's."""
import numpy as np
import pytest

import fdtd_amd as F
from tests.test_gpu_batch import make_case, single

pytestmark = pytest.mark.gpu


def series_equal(a, b):
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


def test_serve_alone_equals_mix():
    probe = make_case(500, (36, 28, 30), 6, 40, nprobe=4)
    others = [make_case(501 + q, s, 5, n, storage=st) for q, (s, n, st) in enumerate([
        ((20, 20, 20), 25, F.STORE_F32), ((48, 16, 24), 33, F.STORE_F16), ((30, 44, 18), 19, F.STORE_F32),
        ((64, 10, 40), 37, F.STORE_BF16), ((17, 23, 29), 50, F.STORE_F32), ((40, 40, 12), 12, F.STORE_F32_FAST)])]
    srv = F.Server(1 << 30, max_slots=8, segment_steps=6)
    print("\n[serve] alone", flush=True)
    r = srv.submit(probe)
    assert srv.wait(r, 60.0)
    alone, info = srv.result(r)
    assert info["state"] == F.REQ_DONE and info["rows"] == probe.steps
    print("[serve] mix", flush=True)
    rs = [srv.submit(x) for x in others[:3]] + [srv.submit(probe)] + [srv.submit(x) for x in others[3:]]
    assert srv.drain(120.0)
    mixed = dict(zip(rs, others[:3] + [probe] + others[3:]))
    for rq, inst in mixed.items():
        got, inf = srv.result(rq)
        assert inf["state"] == F.REQ_DONE, inf
        assert inf["t_submit"] <= inf["t_admit"] <= inf["t_done"]
        want = alone if inst is probe else single(inst)[1]
        assert series_equal(got, want), f"request {rq} {inst.shape}"
    st = srv.stats()
    assert st["completed"] == 8 and st["failed"] == 0 and st["queued"] == 0 and st["running"] == 0
    assert 0 < st["p50_latency_s"] <= st["p99_latency_s"]
    assert st["peak_running"] >= 2   # the mix overlapped on the device
    srv.close()


def test_serve_admits_against_budget():
    insts = [make_case(600 + q, (40, 36, 32), 5, 15 + 3 * q) for q in range(6)]
    one = insts[0].nbytes()
    srv = F.Server(int(2.5 * one), max_slots=16, segment_steps=4)
    print("\n[serve] budget of 2 grids, 6 requests", flush=True)
    rs = [srv.submit(x) for x in insts]
    assert srv.drain(120.0)
    st = srv.stats()
    assert st["completed"] == 6 and st["peak_running"] <= 2 and st["peak_bytes"] <= int(2.5 * one)
    for rq, inst in zip(rs, insts):
        got, inf = srv.result(rq)
        assert inf["state"] == F.REQ_DONE
        assert series_equal(got, single(inst)[1])
    with pytest.raises(F.FdtdError, match="-5"):
        srv.submit(make_case(9, (96, 96, 96), 4, 5))   # larger than the whole budget
    srv.close()


def test_serve_concurrent_clients_and_release():
    """Four client threads submit, wait for and release their own requests
    while the worker runs them together; each result equals a single
    context's.  A running request cannot be released, and a released one is
    unknown afterwards."""
    import threading
    cases = [make_case(700 + q, s, 5, n, storage=st) for q, (s, n, st) in enumerate([
        ((24, 20, 18), 21, F.STORE_F32), ((30, 16, 22), 17, F.STORE_F16), ((18, 26, 20), 29, F.STORE_F32),
        ((20, 20, 30), 13, F.STORE_BF16), ((34, 18, 16), 25, F.STORE_F32), ((22, 22, 22), 19, F.STORE_F32_FAST),
        ((16, 30, 24), 23, F.STORE_F32), ((26, 24, 14), 15, F.STORE_F16)])]
    want = [single(c)[1] for c in cases]
    srv = F.Server(1 << 30, max_slots=6, segment_steps=4)
    print("\n[serve] 4 client threads x 2 requests", flush=True)
    errors = []

    def client(t):
        try:
            for q in (2 * t, 2 * t + 1):
                r = srv.submit(cases[q])
                assert srv.wait(r, 120.0)
                got, inf = srv.result(r)
                assert inf["state"] == F.REQ_DONE, inf
                assert series_equal(got, want[q]), f"case {q}"
                srv.release(r)
        except Exception as e:  # reported by the main thread
            errors.append(f"client {t}: {e!r}")

    threads = [threading.Thread(target=client, args=(t,)) for t in range(4)]
    for th in threads:
        th.start()
    for th in threads:
        th.join(300.0)
    assert not errors, errors
    st = srv.stats()
    assert st["submitted"] == 8 and st["completed"] == 8 and st["failed"] == 0
    long_run = make_case(800, (48, 40, 36), 5, 2000)
    r = srv.submit(long_run)
    with pytest.raises(F.FdtdError, match="-3"):
        srv.release(r)                      # still queued or running
    assert srv.wait(r, 120.0)
    srv.release(r)
    with pytest.raises(F.FdtdError, match="-1"):
        srv.wait(r, 0.0)                    # released: unknown
    srv.close()


def test_serve_failed_admission_leaves_the_rest():
    good = [make_case(900 + q, (32, 28, 24), 5, 20 + q) for q in range(3)]
    bad = make_case(905, (32, 28, 24), 5, 20)
    bad.ids[7] = 5                          # material id == nmat: rejected when the grid is added
    srv = F.Server(1 << 30, max_slots=8, segment_steps=4)
    print("\n[serve] one request with a material id past its table", flush=True)
    rs = [srv.submit(good[0]), srv.submit(bad), srv.submit(good[1]), srv.submit(good[2])]
    assert srv.drain(120.0)
    got, inf = srv.result(rs[1])
    assert inf["state"] == F.REQ_FAILED and inf["error"] == F.ERANGE and inf["rows"] == 0 and got.shape[0] == 0
    for rq, inst in zip(rs[:1] + rs[2:], good):
        series, inf = srv.result(rq)
        assert inf["state"] == F.REQ_DONE
        assert series_equal(series, single(inst)[1])
    st = srv.stats()
    assert st["completed"] == 3 and st["failed"] == 1 and st["running"] == 0
    srv.close()
