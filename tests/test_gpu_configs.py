"""The batched, served and reduced-precision workloads at their real sizes.

  * batched set: all 64 uneven grids (48^3 to 160x160x224, 4-12 Debye
    materials each, a point source, 2 probes) in one batch, autotuned, each
    run for min(its step count, 40) steps.  Every grid's probe series equals
    its single context's bit for bit, the newest fields of every eighth grid
    too, and the largest grid's fields and probes equal the oracle's FP32
    transcription bit for bit.
  * served stream: the first 24 arrivals of the Poisson stream (edges
    32-192, fp32 or fp16 storage, point or 10-cell dipole sources), each
    capped at 200 steps.  Each request's probe series is the same submitted
    alone and in the mix (arriving at its stream time while others run), bit
    for bit; the fp32 requests are within rel-L2 1e-4 of the FP64 oracle.
  * 320x320x880 with fp16 and bf16 storage: the newest fields after 60 steps
    and the probe series against the FP64 oracle, per field, bounded near the
    measured values.

Each case stands in for the reference's step loop (fdtd.cu:115-530) run once
per grid; the single context is itself bit-equal to the oracle's FP32
transcription (test_gpu_parity.py, test_gpu_large.py)."""
import time

import numpy as np
import pytest

import fdtd_amd as F
from fdtd_amd import workloads as W
from oracle import oracle as O
from tests.helpers import bits_equal, first_diff, rel_l2

pytestmark = pytest.mark.gpu


def _oracle(w, prec, steps):
    s = O.Sim(w.shape, w.params.dt, w.params.dx, w.params.dy, w.params.dz, w.params.mu0, w.params.eps0, prec=prec)
    s.set_materials(*w.material_arrays())
    s.set_sources(w.sources, w.jz, jz_mode=1)
    series = s.run(steps, w.probes)
    return s, series


def _single(w, steps, storage=F.STORE_F32):
    ctx = F.Context(w.params, storage=storage)
    ctx.set_materials(w.ids, w.table)
    ctx.set_sources(w.sources, w.jz)
    ctx.set_probes(w.probes, steps)
    ctx.run(steps)
    return ctx


def test_batched_set_all_64_grids():
    cap = 40
    ws = W.batched_set()
    assert len(ws) == 64
    insts = [F.Instance(w.params, F.make_ext(), w.ids, w.table, w.sources, w.jz, min(w.steps, cap), w.probes)
             for w in ws]
    arena = int(sum(x.nbytes() for x in insts) * 1.02) + (64 << 20)
    print(f"\n[configs] batched set: 64 grids, {sum(w.cells for w in ws) / 1e6:.1f} M cells, arena "
          f"{arena / 2**30:.1f} GiB, {cap} steps", flush=True)
    with F.Batch(arena) as b:
        ids = [b.add(x) for x in insts]
        ty, kc = b.autotune(trial_steps=2)
        print(f"[configs] autotuned tile rows {ty}, z planes {kc}", flush=True)
        b.run(cap)
        assert b.info()["running"] == 0
        for q, (i, w, x) in enumerate(zip(ids, ws, insts)):
            assert b.status(i) == (x.steps, x.steps)
            got = b.probes(i)
            with _single(w, x.steps) as ctx:
                want = ctx.probes()
                assert got.shape == want.shape == (x.steps, 2)
                assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), f"grid {q} {w.shape}: probes"
                if q % 8 == 0:
                    for f in range(9):
                        a, c = b.get(i, f), ctx.get(f)
                        assert bits_equal(a, c), f"grid {q} {w.shape} {F.FIELDS[f]}: {first_diff(a, c)}"
        big = max(range(64), key=lambda q: ws[q].cells)
        w = ws[big]
        print(f"[configs] largest grid {w.shape} against the FP32 transcription", flush=True)
        sim, series = _oracle(w, 32, insts[big].steps)
        assert np.array_equal(b.probes(ids[big]).view(np.uint64), series.view(np.uint64))
        assert np.abs(series).max() > 0
        for f in range(9):
            a, c = b.get(ids[big], f), sim.get(f)
            assert bits_equal(a, c), f"largest grid {F.FIELDS[f]}: {first_diff(a, c)}"


def test_served_stream_prefix_alone_and_mixed():
    count, cap = 24, 200
    stream = W.served_stream(count)
    reqs = []
    for t, w, st in stream:
        ext = F.make_ext(storage=F.STORE_F16 if st == "f16" else F.STORE_F32)
        reqs.append((t, w, st, F.Instance(w.params, ext, w.ids, w.table, w.sources, w.jz, min(w.steps, cap),
                                          w.probes)))
    budget = 64 << 30
    print(f"\n[configs] served prefix: {count} requests, {sum(st == 'f16' for _, _, st, _ in reqs)} fp16, "
          f"edges {min(min(w.shape) for _, w, _, _ in reqs)}-{max(max(w.shape) for _, w, _, _ in reqs)}", flush=True)
    alone = []
    with F.Server(budget, max_slots=32, segment_steps=8) as srv:
        for _, _, _, x in reqs:
            r = srv.submit(x)
            assert srv.wait(r, 120.0)
            s, info = srv.result(r)
            assert info["state"] == F.REQ_DONE
            alone.append(s)
            srv.release(r)
    print("[configs] alone done; mix at the stream's arrival times", flush=True)
    with F.Server(budget, max_slots=32, segment_steps=8) as srv:
        t0 = time.perf_counter()
        rs = []
        for t, _, _, x in reqs:
            d = t - (time.perf_counter() - t0)
            if d > 0:
                time.sleep(d)
            rs.append(srv.submit(x))
        assert srv.drain(300.0)
        st = srv.stats()
        print(f"[configs] mix: {st['completed']} solves, {st['solves_per_s']:.1f} solves/s, p50 "
              f"{st['p50_latency_s'] * 1e3:.0f} ms, p99 {st['p99_latency_s'] * 1e3:.0f} ms, peak running "
              f"{st['peak_running']}", flush=True)
        assert st["completed"] == count and st["failed"] == 0 and st["peak_running"] >= 2
        assert 0 < st["p50_latency_s"] <= st["p99_latency_s"]
        for q, (r, (_, w, sto, x)) in enumerate(zip(rs, reqs)):
            got, info = srv.result(r)
            assert info["state"] == F.REQ_DONE and got.shape == (x.steps, 2)
            assert np.array_equal(got.view(np.uint64), alone[q].view(np.uint64)), f"request {q} {w.shape} {sto}"
    O.set_threads(16)
    worst = 0.0
    for q, (_, w, sto, x) in enumerate(reqs):
        if sto != "f32":
            continue
        _, ref = _oracle(w, 64, x.steps)
        for p in range(2):
            if np.linalg.norm(ref[:, p]) > 0:
                e = rel_l2(alone[q][:, p], ref[:, p])
                worst = max(worst, e)
                assert e <= 1e-4, f"request {q} {w.shape} probe {p}: rel-L2 {e:.2e} against FP64"
    print(f"[configs] fp32 requests: worst probe rel-L2 against FP64 {worst:.1e}", flush=True)


# rel-L2 bounds against FP64 after 60 steps, per field (Dx Dy Dz Hx Hy Hz Ex Ey Ez), and on the probe
# beside the dipole, for fp16 and bf16 storage.  Measured on MI355X: fp16 Dx 9.1e-4 Dy 1.2e-3 Dz 1.8e-3
# Hx 5.6e-3 Hy 4.2e-3 Hz 6.8e-2 Ex 6.8e-4 Ey 8.5e-4 Ez 7.7e-4, probe 3.0e-3; bf16 Dx 7.4e-3 Dy 9.1e-3
# Dz 1.2e-2 Hx 4.3e-2 Hy 3.9e-2 Hz 0.64 Ex 6.3e-3 Ey 7.3e-3 Ez 6.8e-3, probe 1.5e-2.  Hz, the smallest
# component beside a z-directed dipole, carries the largest relative error.
BOUNDS_16 = {"f16": ((3e-3, 3e-3, 3e-3, 1e-2, 1e-2, 0.1, 2e-3, 2e-3, 2e-3), 5e-3),
             "bf16": ((2e-2, 2e-2, 2e-2, 0.1, 0.1, 1.0, 2e-2, 2e-2, 2e-2), 3e-2)}


def test_large_phantom_16bit_storage_against_fp64():
    steps = 60
    w = W.large_phantom(steps=steps)
    print(f"\n[configs] large_phantom {w.shape} fp16 and bf16, {steps} steps; FP64 oracle", flush=True)
    O.set_threads(16)
    t0 = time.perf_counter()
    sim, ref = _oracle(w, 64, steps)
    ref_fields = [sim.get(f) for f in range(9)]
    del sim
    print(f"[configs] FP64 oracle {time.perf_counter() - t0:.0f} s", flush=True)
    for name, storage in (("f16", F.STORE_F16), ("bf16", F.STORE_BF16)):
        with _single(w, steps, storage) as ctx:
            errs = [rel_l2(ctx.get(f), ref_fields[f]) for f in range(9)]
            series = ctx.probes()
        near = rel_l2(series[:, 1], ref[:, 1])
        print(f"[configs] {name}: fields " + " ".join(f"{F.FIELDS[f]} {e:.1e}" for f, e in enumerate(errs))
              + f"; probe beside the dipole {near:.1e}", flush=True)
        fb, pb = BOUNDS_16[name]
        for f in range(9):
            assert errs[f] <= fb[f], f"{name} {F.FIELDS[f]}: rel-L2 {errs[f]:.2e} against FP64"
        assert near <= pb, f"{name}: probe beside the dipole rel-L2 {near:.2e}"
