"""Long accuracy runs against the oracle: too long for the test suite, run by
hand on a GPU host (the oracle runs on the host's cores, up to 16 threads).

    python -m tests.long_runs small    # plain_small, 4000 steps: frozen 16-bit error
    python -m tests.long_runs served   # fp16 requests of the served stream vs FP64
    python -m tests.long_runs large    # 320x320x880, 5000 steps, fp32/fp16/bf16 vs FP64
                                       # and the FP32 transcription (about 25 min of host time)
    python -m tests.long_runs large32  # the same, fp32 storage against the transcription only
                                       # (about 11 min of host time)
    python -m tests.long_runs bounds   # peak fields of soft Jz sources against the accumulated
                                       # and the largest kick (the basis of the 16-bit scales;
                                       # oracle only, no GPU)
    python -m tests.long_runs two_step # 320x320x880, 5000 steps: two steps per launch against
                                       # one, every field, level, probe and slice (GPU only)
    python -m tests.long_runs zero_terms  # 320x320x880, 5000 steps: zero-term half rows against
                                          # every row loading, bit for bit (GPU only)
    python -m tests.long_runs batched  # the 64-grid batched set over its full 527-2978 steps in one
                                       # batch against single contexts, bit for bit (GPU only)
    python -m tests.long_runs served_full  # the 120-request served stream at full lengths and its
                                           # arrival times, all fp32 and then half fp16, against
                                           # single contexts (GPU only)
    python -m tests.long_runs riken_body   # the RIKEN phantom at body scale, 200 steps, against the
                                           # FP32 transcription, every field and probe

Each prints one line per case.  The oracle is the checker here, never the
thing measured.
"""
import os
import sys
import time
import warnings

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import fdtd_amd as F  # noqa: E402
from fdtd_amd import workloads as W  # noqa: E402
from oracle import oracle as O  # noqa: E402

STORAGES = (("fp32", F.STORE_F32), ("fp16", F.STORE_F16), ("bf16", F.STORE_BF16))


def rel_l2(x, ref):
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    den = np.linalg.norm(ref)
    return float(np.linalg.norm(x - ref) / den) if den > 0 else float(np.linalg.norm(x - ref))


def oracle_sim(w, prec):
    s = O.Sim(w.shape, w.params.dt, w.params.dx, w.params.dy, w.params.dz, w.params.mu0, w.params.eps0, prec=prec)
    s.set_materials(*w.material_arrays())
    s.set_sources(w.sources, w.jz, jz_mode=1)
    return s


def gpu_ctx(w, storage, nprobe_steps):
    c = F.Context(w.params, storage=storage)
    c.set_materials(w.ids, w.table)
    c.set_sources(w.sources, w.jz)
    c.set_probes(w.probes, nprobe_steps)
    return c


def small():
    """plain_small, 4000 steps: |GPU - FP64| over the largest FP64 norm of
    Dz, Hy, Ez at each checkpoint, and the probe series."""
    steps, checks = 4000, (150, 500, 1000, 2000, 4000)
    w = W.plain_small(steps=steps)
    s = oracle_sim(w, 64)
    ref, series, done = {}, [], 0
    for c in checks:
        series.append(s.run(c - done, w.probes))
        done = c
        ref[c] = {f: s.get(f) for f in (2, 4, 8)}
    series = np.concatenate(series)
    peak = {f: max(np.linalg.norm(ref[c][f]) for c in checks) for f in (2, 4, 8)}
    for name, st in STORAGES[1:]:
        ctx = gpu_ctx(w, st, steps)
        done, row = 0, []
        for c in checks:
            ctx.run(c - done)
            done = c
            row.append(" / ".join(f"{np.linalg.norm(ctx.get(f).astype(np.float64) - ref[c][f]) / peak[f]:.1e}"
                                  for f in (2, 4, 8)))
        print(f"{name} Dz / Hy / Ez at steps {checks}: " + " | ".join(row)
              + f" | Ez probe series {rel_l2(ctx.probes()[:, 0], series[:, 0]):.1e}", flush=True)
        ctx.close()


def served(count=10):
    """The first `count` fp16 requests of the served stream small enough for
    a quick oracle run, each alone in every storage, against FP64."""
    stream = W.served_stream(120)
    picked = [(q, w) for q, (_, w, st) in enumerate(stream)
              if st == "f16" and np.prod(w.shape) * w.steps <= 4e9][:count]
    for q, w in picked:
        ref = oracle_sim(w, 64).run(w.steps, w.probes)
        errs = []
        for name, st in STORAGES:
            ctx = gpu_ctx(w, st, w.steps)
            ctx.run(w.steps)
            p = ctx.probes()
            errs.append(f"{name} {rel_l2(p[:, 0], ref[:, 0]):.1e}/{rel_l2(p[:, 1], ref[:, 1]):.1e}")
            ctx.close()
        dist = np.linalg.norm(w.probes[0, 1:].astype(float) - w.sources[0].astype(float))
        print(f"request {q} {w.shape} {w.steps} steps, {len(w.sources)} source cells, centre probe {dist:.0f} "
              f"cells away; centre/near rel-L2: " + " | ".join(errs), flush=True)


def large(steps=5000, fp64=True):
    """large_phantom over `steps`: probe series and the z = nz/2 planes of Dz,
    Hy, Ez at 1000, 2500 and `steps` against FP64 and the FP32
    transcription (fp64=False: fp32 storage against the transcription
    only)."""
    checks = tuple(c for c in (1000, 2500) if c < steps) + (steps,)
    w = W.large_phantom(steps=steps)
    k = w.shape[2] // 2

    def run_oracle(prec):
        s = oracle_sim(w, prec)
        planes, series, done, t0 = {}, [], 0, time.perf_counter()
        while done < steps:
            n = min(100, steps - done)
            series.append(s.run(n, w.probes))
            done += n
            if done in checks:
                for f in (2, 4, 8):
                    planes[(f, done)] = s.get(f).reshape(w.shape[::-1])[k].copy()
            print(f"  oracle prec {prec}: step {done}, {time.perf_counter() - t0:.0f} s", flush=True)
        return np.concatenate(series), planes

    gpu = {}
    for name, st in (STORAGES if fp64 else STORAGES[:1]):
        ctx = gpu_ctx(w, st, steps)
        planes, done = {}, 0
        for c in checks:
            ctx.run(c - done)
            done = c
            for f in (2, 4, 8):
                planes[(f, c)] = ctx.get(f)[k].copy()
        gpu[name] = (ctx.probes(), planes)
        ctx.close()
    p32, pl32 = run_oracle(32)
    same = np.array_equal(gpu["fp32"][0], p32) and all(
        np.array_equal(gpu["fp32"][1][key].view(np.uint32), pl32[key].view(np.uint32)) for key in pl32)
    print(f"fp32 bit-equal to the FP32 transcription (probes and planes): {same}", flush=True)
    if not fp64:
        return
    p64, pl64 = run_oracle(64)
    for name, (p, planes) in gpu.items():
        probes = " ".join(f"{rel_l2(p[:, q], p64[:, q]):.1e}" for q in range(p.shape[1]))
        pl = " | ".join(f"step {c}: " + " ".join(f"{rel_l2(planes[(f, c)], pl64[(f, c)]):.1e}" for f in (2, 4, 8))
                        for c in checks)
        print(f"{name} vs FP64: probes {probes}; plane Dz Hy Ez {pl}", flush=True)


def _same_runs(w, steps, variants):
    """Runs w over `steps` in each of two context variants (keyword
    arguments of F.Context) and reports whether every field of every ring
    level, the probe series and the slices agree bit for bit."""
    out = []
    for kw in variants:
        t0 = time.perf_counter()
        ctx = F.Context(w.params, **kw)
        ctx.set_materials(w.ids, w.table)
        ctx.set_sources(w.sources, w.jz)
        ctx.set_probes(w.probes, steps)
        ctx.set_slices(w.slice_k, w.slice_every, steps // w.slice_every)
        ctx.set_timing(1)
        ctx.run(steps)
        fields = {(f, lv): ctx.get(f, lv) for f in range(9) for lv in ([0] if 3 <= f < 6 else [0, 1, 2])}
        out.append((fields, ctx.probes(), ctx.slices(), ctx.timed_steps()))
        ctx.close()
        print(f"  {kw}: {steps} steps, {time.perf_counter() - t0:.1f} s", flush=True)
    (fa, pa, sa, ka), (fb, pb, sb, kb) = out
    same_fields = all(np.array_equal(fa[key].view(np.uint32), fb[key].view(np.uint32)) for key in fa)
    same_probes = np.array_equal(pa.view(np.uint64), pb.view(np.uint64))
    same_slices = list(sa[0]) == list(sb[0]) and np.array_equal(sa[1].view(np.uint32), sb[1].view(np.uint32))
    return dict(fields=same_fields, probes=same_probes, slices=same_slices, nslices=len(sa[0]),
                steps_per_launch=(ka, kb), peak=float(np.abs(pa).max()), finite=bool(np.isfinite(pa).all()))


def two_step(steps=5000):
    """large_phantom over `steps` with two steps per launch (time_steps = 2)
    and with one: every field of every ring level, the probe series and the
    slices must agree bit for bit."""
    r = _same_runs(W.large_phantom(steps=steps), steps, [dict(time_steps=2), dict(time_steps=1)])
    ka, kb = r["steps_per_launch"]
    print(f"two-step launches ({ka} steps per timed launch) against one ({kb}): fields {r['fields']}, "
          f"probes {r['probes']}, slices {r['slices']} ({r['nslices']} slices); probe peak "
          f"{r['peak']:.3e}, finite {r['finite']}", flush=True)


def zero_terms(steps=5000):
    """large_phantom over `steps` with zero-term half rows (the default) and
    with every row loading E^{n-1} and D^{n-1} (load_zero_terms = 1): every
    field of every ring level, the probe series and the slices must agree
    bit for bit, through the late field's small and -0 values."""
    r = _same_runs(W.large_phantom(steps=steps), steps, [dict(load_zero_terms=0), dict(load_zero_terms=1)])
    print(f"zero-term half rows against loading every row: fields {r['fields']}, probes {r['probes']}, "
          f"slices {r['slices']} ({r['nslices']} slices); probe peak {r['peak']:.3e}, finite {r['finite']}",
          flush=True)


def batched():
    """All 64 grids of the batched set over their full step counts
    (527-2978) in one batch (autotuned shape, graphs, runs of tiles per
    XCD), against each grid stepped alone in a fused context: every probe
    sample and every field after the last step, bit for bit."""
    ws = W.batched_set()
    insts = [F.Instance(w.params, F.make_ext(), w.ids, w.table, w.sources, w.jz, w.steps, w.probes) for w in ws]
    arena = int(sum(x.nbytes() for x in insts) * 1.02) + (64 << 20)
    t0 = time.perf_counter()
    with F.Batch(arena) as b:
        ids = [b.add(x) for x in insts]
        ty, kc = b.autotune(trial_steps=2)
        b.run(max(w.steps for w in ws))
        assert b.info()["running"] == 0
        got = [(b.probes(i), [b.get(i, f) for f in range(9)]) for i in ids]
    print(f"  batch: 64 grids, {ty} tile rows, {kc} planes, {time.perf_counter() - t0:.1f} s", flush=True)
    bad = []
    for q, (w, (pr, fields)) in enumerate(zip(ws, got)):
        ctx = F.Context(w.params)
        ctx.set_materials(w.ids, w.table)
        ctx.set_sources(w.sources, w.jz)
        ctx.set_probes(w.probes, w.steps)
        ctx.run(w.steps)
        same = np.array_equal(pr.view(np.uint64), ctx.probes().view(np.uint64)) and pr.shape[0] == w.steps
        same = same and all(np.array_equal(np.ascontiguousarray(fields[f], np.float32).view(np.uint32),
                                           np.ascontiguousarray(ctx.get(f), np.float32).view(np.uint32))
                            for f in range(9))
        ctx.close()
        if not same:
            bad.append((q, w.shape))
    print(f"batched set, full horizons, batch against single contexts: {64 - len(bad)} of 64 grids bit-equal "
          f"(probes and all nine fields); differing: {bad}", flush=True)


def served_full(count=120):
    """The served stream at every request's full step count (200-4000),
    submitted at its arrival time to one server: all fp32, then with about
    half the requests in fp16 storage (the default stream).  Each request's
    probe series against the same grid stepped alone in a fused context of
    its storage, bit for bit."""
    for share in (0.0, 0.5):
        stream = W.served_stream(count, f16_share=share)
        store = [F.STORE_F16 if st == "f16" else F.STORE_F32 for _, _, st in stream]
        insts = [F.Instance(w.params, F.make_ext(storage=sto), w.ids, w.table, w.sources, w.jz, w.steps, w.probes)
                 for (_, w, _), sto in zip(stream, store)]
        with F.Server(200 << 30, max_slots=64, segment_steps=8) as srv:
            t0 = time.perf_counter()
            rs = []
            for (t, _, _), x in zip(stream, insts):
                d = t - (time.perf_counter() - t0)
                if d > 0:
                    time.sleep(d)
                rs.append(srv.submit(x))
            assert srv.drain(600.0)
            st = srv.stats()
            got = [srv.result(r) for r in rs]
        print(f"  server: {st['completed']} solves, {st['solves_per_s']:.1f} solves/s, peak running "
              f"{st['peak_running']}", flush=True)
        bad = []
        for q, ((_, w, _), sto, (series, info)) in enumerate(zip(stream, store, got)):
            ctx = F.Context(w.params, storage=sto)
            ctx.set_materials(w.ids, w.table)
            ctx.set_sources(w.sources, w.jz)
            ctx.set_probes(w.probes, w.steps)
            ctx.run(w.steps)
            with warnings.catch_warnings():
                warnings.simplefilter("ignore", F.PrecisionWarning)   # compared, not judged, here
                want = ctx.probes()
            ctx.close()
            if info["state"] != F.REQ_DONE or series.shape != want.shape or \
                    not np.array_equal(series.view(np.uint64), want.view(np.uint64)):
                bad.append((q, w.shape))
        nf16 = sum(sto == F.STORE_F16 for sto in store)
        print(f"served stream, {count} requests ({nf16} fp16) at full length, server against single contexts: "
              f"{count - len(bad)} of {count} probe series bit-equal; differing: {bad}", flush=True)


def riken_body(steps=200):
    """The RIKEN phantom at body scale (297x522x1719 with its vacuum padding,
    266.5 M cells) over `steps`: the fused context at its defaults against
    the oracle's FP32 transcription, every probe sample and every field, bit
    for bit."""
    from fdtd_amd import phantom as P
    table, _ = F.load_tissue_table(os.path.join(ROOT, "tests", "golden", "mat_specs_riken"))
    w = W.riken(P.load(os.path.join(ROOT, "data", "riken_ids.npz"), P.RIKEN_SHAPE), table, steps=steps)
    ctx = gpu_ctx(w, F.STORE_F32, steps)
    ctx.run(steps)
    s = oracle_sim(w, 32)
    series, done, t0 = [], 0, time.perf_counter()
    while done < steps:
        n = min(20, steps - done)
        series.append(s.run(n, w.probes))
        done += n
        print(f"  oracle: step {done}, {time.perf_counter() - t0:.0f} s", flush=True)
    ref = np.concatenate(series)
    same_probes = np.array_equal(ctx.probes().view(np.uint64), ref.view(np.uint64))
    differ = [F.FIELDS[f] for f in range(9) if not np.array_equal(ctx.get(f).view(np.uint32), s.get(f).view(np.uint32))]
    peak = float(np.abs(ref).max())
    ctx.close()
    print(f"riken {w.shape}, {steps} steps, GPU against the FP32 transcription: probes {same_probes} (peak {peak:.3e}), "
          f"fields differing: {differ}", flush=True)


def bounds(n=40):
    """FP64 oracle in an n^3 box: the largest |E|, |D|/eps0 and eta0 |H| a
    soft Jz source drives, over K = max_t |sum jz| / eps0 (the accumulated
    kick) and over the largest single kick max |jz| / eps0, for point, 10-cell
    line and full-plane sources at 0.3-6 GHz, in vacuum and at eps_r = 50."""
    c = n // 2
    shapes = {"point": [[c, c, c]], "line10": [[c, c, z] for z in range(c - 5, c + 5)],
              "sheet": [[10, y, z] for y in range(1, n - 1) for z in range(0, n - 1)]}
    cases = [(f, k, 1.0) for f in (0.3e9, 0.9e9, 3e9, 6e9) for k in ("point", "line10")]
    cases += [(0.9e9, "sheet", 1.0), (3e9, "sheet", 1.0), (0.9e9, "point", 50.0)]
    for freq, kind, eps in cases:
        steps = 1200 if freq < 0.5e9 else 600
        params, jz = W.derive((n, n, n), steps, freq, 2.0e-3, 0.99)
        eta0 = np.sqrt(params.mu0 / params.eps0)
        kick = np.abs(jz).max() / params.eps0
        K = np.abs(np.cumsum(jz.astype(np.float64))).max() / params.eps0
        s = O.Sim((n, n, n), params.dt, params.dx, params.dy, params.dz, params.mu0, params.eps0, prec=64)
        z = np.zeros((n, n, n), np.float32)
        s.set_materials(z, z + eps, z + eps, z)
        s.set_sources(np.asarray(shapes[kind], np.int32), jz, jz_mode=1)
        last = int(np.flatnonzero(jz).max()) if np.any(jz) else 0
        mE = mD = mH = 0.0
        for _ in range(0, min(steps, last + 100), 10):
            s.step(10)
            mE = max(mE, max(np.abs(s.get(f)).max() for f in (6, 7, 8)))
            mD = max(mD, max(np.abs(s.get(f)).max() for f in (0, 1, 2)) / params.eps0)
            mH = max(mH, max(np.abs(s.get(f)).max() for f in (3, 4, 5)) * eta0)
        print(f"{freq / 1e9:4.1f} GHz {kind:6s} eps_r {eps:4.1f}: max E/K {mE / K:.3f}, D/eps0/K {mD / K:.3f}, "
              f"eta0 H/kick {mH / kick:.3f}", flush=True)


if __name__ == "__main__":
    O.set_threads(min(os.cpu_count() or 1, 16))
    {"small": small, "served": served, "large": large, "bounds": bounds,
     "two_step": two_step, "zero_terms": zero_terms, "batched": batched,
     "served_full": served_full, "riken_body": riken_body, "large32": lambda: large(fp64=False)}[sys.argv[1] if len(sys.argv) > 1 else "small"]()
