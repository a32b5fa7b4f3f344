"""Benchmark of the FTLM hot path on one MI355X.

Workload: the largest block of the 4x4 sweep (lx=ly=4, t=1, t'=0, U=7, V=0):
(N_up, N_down, k) = (8, 8, (0, 0)), dim = 10,353,252, with all R = 128
seeded random vectors advancing together for 80 Lanczos steps. One bench
step is what the drop-in executable runs for that block (ftlm_ftlm_block on
its default path): the seeded start vectors drawn on the GPU (the reference's
mt19937_64 stream), 80 batched H applies + tree-order recurrences, and the
QL Ritz values and weights copied back. The block's tables and hop list are
resident in HBM before the timed region.

value = algorithmic H-apply bandwidth: sum over applies of
        dim * R * 16 B * (hbar + 2)  /  summed H-apply kernel time
(hbar = distinct non-zero hop targets per row, 34.13 here), measured with HIP
events on the engine's stream. Beside it: the plain per-vector GPU path (the
same engine with R = 1 kernels) on a few of the same vectors, and the
reference CPU code (oracle/_ref, the reference's run_ftlm_block compiled from
its sources) timed on a bounded block of the same sweep.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

METRIC = "4x4 sweep sec: batched vs plain GPU FTLM at |dn|<=1e-10; H-apply HBM GB/s"
HBM_PEAK_GBS = 8000.0


def hbar(L, nup, ndn, d):
    return d * (nup * (L - nup) + ndn * (L - ndn)) / (L - 1)


def cpu_baseline(budget_s=25.0):
    """Reference run_ftlm_block (compiled from the reference sources into
    oracle/_ref) on a bounded block of the same 4x4 sweep; reported in the
    same unit (algorithmic H-apply GB/s). Falls back to the oracle restatement
    when the harness was not built."""
    import numpy as np  # noqa: F401
    import oracle_lib

    kind = "reference"
    lib = oracle_lib.reference()
    if lib is None:
        lib = oracle_lib.oracle()
        kind = "port"
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    cpu_model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu_model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    model = oracle_lib.model(lx=4, ly=4, u=7.0)
    # a mid-size block of the same sweep at the workload's R and steps
    # (~20 s on 16 cores; the reference's serial basis build is not timed)
    nup, ndn, mx, my = 4, 4, 0, 0
    blk = lib.block(model, nup, ndn, mx, my)
    steps = 80
    R = 128
    bs = lib.block_seed(12345, nup, ndn, mx, my)
    t0 = time.perf_counter()
    lib.ftlm_block(blk, R, steps, bs, threads=threads)
    dt = time.perf_counter() - t0
    m = min(steps, blk.dim)
    byts = m * blk.dim * R * 16.0 * (hbar(16, nup, ndn, 4) + 2.0)
    return {"value": round(byts / dt / 1e9, 3), "unit": "GB/s", "cores": threads, "kind": kind, "cpu": cpu_model,
            "sample": (f"4x4 U=7 block (Nup={nup},Ndown={ndn},k=({mx},{my})) dim={blk.dim}, R={R}, {steps} steps, "
                       f"run_ftlm_block incl. Jacobi, {dt:.1f} s on {threads} threads")}


def _interp_log(xs, ys, x):
    """Piecewise-linear interpolation in log(x), clamped at the ends."""
    import numpy as np

    return float(np.interp(np.log(x), np.log(np.asarray(xs, float)), np.asarray(ys, float)))


def sweep_estimate(ftlm, R=128, m=80, pack_every=8, validate=0, log=None):
    """Seconds of the whole 4x4 U=7 sweep (the CLI's block loop over all 4,564
    (N_up, N_down, k) blocks, R vectors, m Lanczos steps), batched and plain
    per-vector, from measured parts:

    * basis: every sector and block built (measured, whole sweep);
    * exact blocks (dim <= 256): all 708 in one queued eigensolve (measured);
    * packed FTLM blocks (256 < dim <= 131072): every `pack_every`-th of them
      in one persistent launch, extrapolated by algorithmic bytes;
    * solo FTLM blocks (dim > 131072): a stratified sample by dimension
      (quantiles of the solo range, the largest block included), each run
      batched; seconds per byte interpolated in log(dim) over all solo blocks
      (--sweep-validate N times N other solo blocks against that prediction);
    * plain per-vector mode: one vector of sampled blocks over the whole
      dimension range with the R = 1 kernels, seconds per vector interpolated
      in log(dim), times R per block.
    Device time of the Lanczos passes; the start vectors are drawn on the GPU
    (the CLI's default with tree reductions), their time reported separately."""
    import numpy as np

    say = log or (lambda s: None)
    L, lx, ly = 16, 4, 4
    eng = ftlm.Engine(lx=lx, ly=ly, u=7.0)
    eng.set_start_source(ftlm.START_DEVICE)  # the CLI's default with tree reductions
    t0 = time.perf_counter()
    blocks = []
    for nup in range(L + 1):
        for ndn in range(L + 1):
            eng.sector(nup, ndn)
            for my in range(ly):
                for mx in range(lx):
                    d = eng.block(mx, my)
                    if d:
                        blocks.append((nup, ndn, mx, my, d))
    t_basis = time.perf_counter() - t0
    say(f"# sweep: {len(blocks)} blocks enumerated in {t_basis:.2f}s")

    def algo_bytes(b):
        nup, ndn, _, _, d = b
        k = min(m, d)
        return k * d * R * 16.0 * (hbar(L, nup, ndn, 4) + 2.0) + k * 4.0 * d * R * 16.0  # apply + recurrence

    exact = [b for b in blocks if b[4] <= 256]
    packed = [b for b in blocks if 256 < b[4] <= 131072]
    solo = [b for b in blocks if b[4] > 131072]

    # exact blocks: all, one queued eigensolve
    t0 = time.perf_counter()
    eng.exact_queue_begin()
    cur = None
    for b in exact:
        if cur != b[:2]:
            eng.sector(*b[:2])
            cur = b[:2]
        eng.exact_queue_add(b[2], b[3])
    eng.exact_queue_run()
    t_exact = time.perf_counter() - t0
    say(f"# sweep: {len(exact)} exact blocks {t_exact:.2f}s")

    # packed blocks: a deterministic 1/pack_every sample in one pack
    sample = packed[::pack_every]
    t0 = time.perf_counter()
    eng.pack_begin(R, m)
    cur = None
    for b in sample:
        if cur != b[:2]:
            eng.sector(*b[:2])
            cur = b[:2]
        eng.block(b[2], b[3])
        if eng.pack_add(ftlm.block_seed(12345, *b[:4])) is None:
            raise RuntimeError("pack sample does not fit one pack")
    t_fill = time.perf_counter() - t0
    eng.pack_tune(3)
    st = eng.pack_run(0, 0, 1)
    pack_s = st["ms_total"] / 1000.0
    pack_rate = sum(algo_bytes(b) for b in sample) / pack_s
    t_packed = sum(algo_bytes(b) for b in packed) / pack_rate
    say(f"# sweep: pack of {len(sample)}/{len(packed)} blocks {pack_s:.3f}s device, fill {t_fill:.1f}s "
        f"-> {t_packed:.1f}s for all packed blocks")

    # solo blocks and plain per-vector: stratified samples
    dsolo = sorted(b[4] for b in solo)
    want = [dsolo[int(q * (len(dsolo) - 1))] for q in (0.0, 0.25, 0.5, 0.75, 0.9)] + [dsolo[-1]]
    pick = []
    for w in want:
        b = min(solo, key=lambda x: (abs(x[4] - w), x))
        if b not in pick:
            pick.append(b)
    dpk = sorted(b[4] for b in packed)
    plain_pick = [min(packed, key=lambda x: (abs(x[4] - dpk[int(q * (len(dpk) - 1))]), x))
                  for q in (0.0, 0.33, 0.66, 1.0)]
    solo_rate, plain_vec = [], []
    t_host_starts = 0.0
    # one untimed plain pass first: the first R = 1 launch pays one-time costs
    b = plain_pick[0]
    eng.sector(*b[:2])
    eng.block(b[2], b[3])
    eng.prepare(R, ftlm.block_seed(12345, *b[:4]))
    eng.run(R, m, ftlm.REDUCE_TREE, per_vector_column=0)
    for b in plain_pick + pick:
        eng.sector(*b[:2])
        d = eng.block(b[2], b[3])
        a = time.perf_counter()
        eng.prepare(R, ftlm.block_seed(12345, *b[:4]))
        t_host_starts += time.perf_counter() - a
        if b in pick:
            s = eng.run(R, m, ftlm.REDUCE_TREE)
            a = time.perf_counter()
            eng.finish(R, m)
            fin = time.perf_counter() - a
            secs = s["ms_total"] / 1000.0 + fin
            solo_rate.append((d, algo_bytes(b) / secs))
            say(f"# sweep: solo block {b[:4]} dim {d}: {secs:.3f}s ({algo_bytes(b) / secs / 1e12:.2f} TB/s)")
        a = time.perf_counter()
        eng.run(R, m, ftlm.REDUCE_TREE, per_vector_column=0)
        pv = time.perf_counter() - a
        plain_vec.append((d, pv))
        say(f"# sweep: plain vector {b[:4]} dim {d}: {pv * 1000:.2f} ms")
    solo_rate.sort()
    plain_vec.sort()
    # check of the interpolation on solo blocks outside the sample
    checks = []
    if validate > 0:
        import numpy as np

        rng = np.random.default_rng(7)
        others = [b for b in solo if b not in pick]
        for j in rng.choice(len(others), size=min(validate, len(others)), replace=False):
            b = others[int(j)]
            eng.sector(*b[:2])
            d = eng.block(b[2], b[3])
            eng.prepare(R, ftlm.block_seed(12345, *b[:4]))
            s = eng.run(R, m, ftlm.REDUCE_TREE)
            a = time.perf_counter()
            eng.finish(R, m)
            secs = s["ms_total"] / 1000.0 + time.perf_counter() - a
            pred = algo_bytes(b) / _interp_log([x[0] for x in solo_rate], [x[1] for x in solo_rate], d)
            a = time.perf_counter()
            eng.run(R, m, ftlm.REDUCE_TREE, per_vector_column=0)
            pv = time.perf_counter() - a
            pv_pred = _interp_log([x[0] for x in plain_vec], [x[1] for x in plain_vec], d)
            checks.append({"block": list(b[:4]), "dim": d, "batched_s": round(secs, 3),
                           "predicted_s": round(pred, 3), "plain_ms": round(pv * 1000, 2),
                           "plain_predicted_ms": round(pv_pred * 1000, 2)})
            say(f"# sweep check {b[:4]} dim {d}: batched {secs:.3f}s vs predicted {pred:.3f}s; "
                f"plain {pv * 1000:.1f} ms vs predicted {pv_pred * 1000:.1f} ms")
    eng.close()
    t_solo = sum(algo_bytes(b) / _interp_log([x[0] for x in solo_rate], [x[1] for x in solo_rate], b[4])
                 for b in solo)
    t_plain_ftlm = sum(R * _interp_log([x[0] for x in plain_vec], [x[1] for x in plain_vec], b[4])
                       for b in packed + solo)
    batched = t_basis + t_exact + t_packed + t_solo
    plain = t_basis + t_exact + t_plain_ftlm
    return {"metric": "4x4 sweep sec (extrapolated from a stratified sample)", "batched_s": round(batched, 1),
            "plain_s": round(plain, 1), "batched_vs_plain": round(plain / batched, 2),
            "parts_s": {"basis": round(t_basis, 2), "exact": round(t_exact, 2), "packed": round(t_packed, 1),
                        "solo": round(t_solo, 1), "plain_ftlm": round(t_plain_ftlm, 1)},
            "blocks": {"exact": len(exact), "packed": len(packed), "solo": len(solo),
                       "packed_sampled": len(sample), "solo_sampled": len(pick),
                       "plain_sampled": len(plain_pick) + len(pick)},
            "solo_rate_TBps": [(d, round(r / 1e12, 3)) for d, r in solo_rate],
            "plain_ms_per_vector": [(d, round(t * 1000, 2)) for d, t in plain_vec],
            "start_vectors_s_sampled": round(t_host_starts, 2), "checks": checks}


def memory_ceiling(eng, R, rows=10353252, k=35, reps=3):
    """HBM rates of the apply's access shapes without its hop metadata
    (ftlm_memprobe): stream read and copy of rows x R complex doubles, and a
    gather writing each row as the sum of k hashed rows (k = 35 ~ the 34.13
    hop targets + the row itself of the (8,8) block, one wavefront per row
    like k_apply_hl), the same gather inside each row's 12,870-row window
    (one canonical-up group: where the down-spin hops land), and half of it
    inside the window (the apply's mix of down- and up-spin hops)."""
    out = {"rows": rows, "width": R, "row_bytes": R * 16}
    for kind, name in ((0, "read"), (1, "copy"), (2, "gather"), (3, "gather_window"), (4, "gather_half_window")):
        ms, by = eng.memprobe(kind, rows, R, k, reps)
        out[name] = {"ms": round(ms, 3), "GB": round(by / 1e9, 2), "TBps": round(by / ms / 1e9, 3)}
        print(f"# ceiling {name}: {ms:.2f} ms, {by / ms / 1e9:.2f} TB/s", file=sys.stderr, flush=True)
    return out


# FLOPs of one surrogate training step at batch B, as the kernels execute
# it (features padded to 32, outputs to 32): forward 2B(32*256 + 3*256^2 +
# 256*32), backward the data gradients of layers 4..1 and the weight
# gradients of layers 4..0 (the same products again, minus the first layer's
# data gradient).
def surrogate_step_flop(B):
    fwd = 2 * B * (32 * 256 + 3 * 256 * 256 + 256 * 32)
    dgrad = 2 * B * (256 * 32 + 3 * 256 * 256)
    wgrad = 2 * B * (32 * 256 + 3 * 256 * 256 + 256 * 32)
    return fwd + dgrad + wgrad


def surrogate_step(steps, eager, B=4096, n=200000):
    """Training steps of the surrogate MLP (bf16 MFMA kernels, fp32
    accumulation) at batch B on synthetic rows over its parameter box:
    replayed HIP graphs (what training runs), or eager launches (`eager`:
    for kernel-trace profiling, which does not follow graph replays)."""
    import numpy as np

    from ftlm_extended_hubbard_development_amd import surrogate

    rng = np.random.default_rng(0)
    lo = np.array([4.0, 0.0, -0.3, 0.0, 0.0, 2.0, -5.0])
    hi = np.array([10.0, 2.0, 0.3, 1.0, 1.0, 20.0, 15.0])
    x = lo + (hi - lo) * rng.random((n, 7))
    y = np.stack([np.tanh((x[:, 6] - x[:, 0] / 2) * x[:, 5] / 10) + 1, np.exp(-((x[:, 6] - 3) / 4) ** 2)], axis=1)
    s = surrogate.Surrogate(y_mean=(1.0, 0.5), y_scale=(1.0, 0.5), batch=B, lr=1e-3, seed=1)
    try:
        if eager:
            xb, yb = x[:B], y[:B]
            s.step(xb, yb)
            t0 = time.perf_counter()
            for _ in range(steps):
                s.step(xb, yb)
            dt = time.perf_counter() - t0
        else:
            s.set_data(x, y)
            s.train(20, seed=1)
            t0 = time.perf_counter()
            s.train(steps, seed=2)
            dt = time.perf_counter() - t0
    finally:
        s.close()
    ms = dt * 1000.0 / steps
    flop = surrogate_step_flop(B)
    return {"mode": "eager" if eager else "graph", "batch": B, "steps": steps, "ms_per_step": round(ms, 4),
            "GFLOP_per_step": round(flop / 1e9, 3), "TFLOPs": round(flop / ms / 1e9, 2),
            "frac_bf16_dense_peak": round(flop / ms / 1e9 / 2500.0, 5)}


def read_measured_sweep():
    """The whole 4x4 sweep as run by the drop-in executable (resumable
    chunks, committed summary): total wall seconds and its parts."""
    path = os.path.join(ROOT, "profiles", "sweep_4x4_measured.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def read_traffic():
    """HBM bytes per H-apply launch from the committed rocprofv3 PMC summary
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration)."""
    path = os.path.join(ROOT, "profiles", "apply_4x4_8_8_R128_pmc.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("hbm_bytes_per_apply")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--R", type=int, default=128)
    ap.add_argument("--lanczos-steps", type=int, default=80)
    ap.add_argument("--block", default="8,8,0,0")
    ap.add_argument("--plain-vectors", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-sweep", action="store_true",
                    help="skip the whole-sweep estimate (batched vs plain, from a stratified block sample)")
    ap.add_argument("--sweep-validate", type=int, default=0, metavar="N",
                    help="also time N random solo blocks outside the sample against the estimate's prediction")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="first run one plain device copy of BYTES bytes (the reference kernel for scaling "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE counters to HBM bytes)")
    ap.add_argument("--ceiling", action="store_true",
                    help="only measure what the HBM path delivers for the apply's shapes (stream read, copy, random "
                         "2 KB-row gather of 35 rows per row) on the workload block's vector size, and print it")
    ap.add_argument("--surrogate", type=int, default=0, metavar="STEPS",
                    help="only time STEPS surrogate training steps (batch 4096, replayed HIP graphs) and print them")
    ap.add_argument("--surrogate-eager", action="store_true",
                    help="with --surrogate: eager launches instead of graph replays (for kernel traces)")
    ap.add_argument("--vector-placement", type=int, default=None, metavar="N",
                    help="buffers the engine times for the Lanczos vectors when it allocates them "
                         "(default: the engine's, 8; 0 off)")
    ap.add_argument("--tile", default="auto",
                    help="H-apply configuration: auto (autotuner) or PATH,W (path 1 hop list / 2 on the fly; "
                         "W -1 automatic, 0 untiled, >0 tile width)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this engine runs each FTLM block on a single device "
              f"(--gpus {args.gpus} is not supported)", file=sys.stderr)
        sys.exit(2)

    import ftlm_extended_hubbard_development_amd as ftlm

    if args.surrogate:
        print(json.dumps(surrogate_step(args.surrogate, args.surrogate_eager)), flush=True)
        return

    if args.copy_calibration:
        import torch

        n = args.copy_calibration // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(f"# copy calibration: {n * 8} bytes read, {n * 8} bytes written", file=sys.stderr, flush=True)
        del src, dst
        torch.cuda.empty_cache()

    nup, ndn, mx, my = (int(t) for t in args.block.split(","))
    eng = ftlm.Engine(lx=4, ly=4, u=7.0)
    if args.vector_placement is not None:
        eng.set_vector_placement(args.vector_placement)
    if args.ceiling:
        print(json.dumps(memory_ceiling(eng, args.R)), flush=True)
        eng.close()
        return
    t0 = time.perf_counter()
    eng.sector(nup, ndn)
    dim = eng.block(mx, my)
    t_basis = time.perf_counter() - t0
    print(f"# block ({nup},{ndn},{mx},{my}) dim={dim} basis {t_basis:.2f}s", file=sys.stderr, flush=True)
    R, m = args.R, args.lanczos_steps
    bseed = ftlm.block_seed(12345, nup, ndn, mx, my)
    # the executable's default: start vectors drawn on the GPU (tree mode)
    eng.set_start_source(ftlm.START_DEVICE)
    t0 = time.perf_counter()
    eng.prepare(R, bseed)
    print(f"# start vectors for tuning ({time.perf_counter() - t0:.2f}s)", file=sys.stderr, flush=True)
    if args.tile == "auto":
        tile, tile_ms, tune = eng.autotune(R)
        print(f"# autotune: {tune} -> {tile}", file=sys.stderr, flush=True)
    else:
        path, w = (int(t) for t in args.tile.split(","))
        eng.set_apply_config(path, w)
        tile, tune = eng.apply_config(R), None

    for w in range(args.warmup):
        eng.ftlm_block(R, m, bseed, reduce=ftlm.REDUCE_TREE)
        print(f"# warmup {w} done", file=sys.stderr, flush=True)
    ms_apply = 0.0
    bytes_apply = 0.0
    n_apply = 0
    t0 = time.perf_counter()
    for k in range(args.steps):
        t1 = time.perf_counter()
        eig, wts = eng.ftlm_block(R, m, bseed, reduce=ftlm.REDUCE_TREE)
        dt = time.perf_counter() - t1
        st = eng.last_stats()
        ms_apply += st["ms_apply"]
        bytes_apply += st["bytes_apply"]
        n_apply += st["n_apply"]
        assert abs(wts.sum() - 1.0) < 1e-10, wts.sum()
        print(f"# step {k}: apply {st['ms_apply'] / st['n_apply']:.2f} ms/launch, Lanczos {st['ms_total']:.0f} ms, "
              f"block {dt * 1000:.0f} ms", file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1000.0 / args.steps
    achieved = bytes_apply / (ms_apply / 1000.0) / 1e9
    place_ms, place_pick = eng.vector_placement_report()
    print(f"# vector placement: apply ms per candidate buffer {[round(float(v), 2) for v in place_ms]}, "
          f"kept {place_pick}", file=sys.stderr, flush=True)

    # plain per-vector GPU FTLM on the same block (R = 1 kernels)
    eng.prepare(R, bseed)
    plain_ms = []
    for c in range(args.plain_vectors):
        t1 = time.perf_counter()
        eng.run(R, m, ftlm.REDUCE_TREE, per_vector_column=c)
        plain_ms.append((time.perf_counter() - t1) * 1000.0)
        print(f"# plain vector {c}: {plain_ms[-1]:.0f} ms", file=sys.stderr, flush=True)
    plain_per_vec = sorted(plain_ms)[len(plain_ms) // 2] if plain_ms else None
    speedup = (plain_per_vec * R / ms_per_step) if plain_per_vec else None
    eng.close()

    sweep = None
    if not args.no_sweep:
        try:
            sweep = sweep_estimate(ftlm, R, m, validate=args.sweep_validate,
                                   log=lambda t: print(t, file=sys.stderr, flush=True))
        except Exception as e:  # reported, not hidden: the per-block measurement above stands
            sweep = {"error": f"{type(e).__name__}: {e}"}
    measured = read_measured_sweep()
    if measured and isinstance(sweep, dict) and "plain_s" in sweep:
        # the batched half of the metric measured end to end through the
        # executable (committed run); the plain per-vector half stays the
        # stratified extrapolation of this run
        sweep["batched_s_measured"] = measured["batched_s"]
        sweep["batched_vs_plain_measured"] = round(sweep["plain_s"] / measured["batched_s"], 2)
        sweep["measured_run"] = {k: measured[k] for k in ("command", "chunks", "wall_s", "blocks") if k in measured}
    cpu = None if args.no_cpu else cpu_baseline()
    hb = hbar(16, nup, ndn, 4)
    per_launch = dim * R * 16.0 * (hb + 2.0)
    out = {
        "metric": METRIC,
        "value": round(achieved, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded unit-norm complex Gaussian start vectors (reference stream, seed 12345)",
        "config": {"workload": f"4x4 U=7 FTLM block (Nup={nup},Ndown={ndn},k=({mx},{my})) dim={dim}, "
                               f"R={R} batched vectors, {m} Lanczos steps, the executable's default path "
                               f"(GPU start vectors, tree reductions, QL Ritz)",
                   "lattice": "4x4", "u": 7.0, "R": R, "lanczos_steps": m, "block_dim": dim,
                   "apply_config": {"path": ["auto", "hoplist", "onthefly"][tile[0]], "tile": tile[1]},
                   "autotune_ms_per_apply": tune,
                   "vector_placement": {"apply_ms_per_candidate": [round(float(v), 2) for v in place_ms],
                                        "kept": place_pick}},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": read_traffic(),
                     "algorithmic_bytes_per_launch": per_launch,
                     "ms_per_launch": round(ms_apply / n_apply, 3)},
        "plain_per_vector_ms": round(plain_per_vec, 2) if plain_per_vec else None,
        "batched_vs_plain": round(speedup, 2) if speedup else None,
        "cpu_baseline": cpu,
        "sweep": sweep,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
