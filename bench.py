#!/usr/bin/env python3
"""Walk-throughput benchmark of the MI355X walk engine (one GPU).

Default workload (Node2vec out of core, learned vs plain block scheduler):
  undirected RMAT graph, 2^26 ids (a,b,c,d = .57,.19,.19,.05), symmetrized, deduplicated,
  self-loops dropped, ids randomly permuted; 1.856B edges are drawn so that at least the
  stated 3.6B arcs remain after deduplication (generated on the GPU, seed 42); cut into
  256 MB blocks; HBM block cache capped at 16 GB (the stated size, where the whole CSR
  fits) and at 8 GB (where blocks must be evicted); p 0.5, q 2, one walker per vertex,
  length 40.  One timed step = one full walk of every walker through the out-of-core
  engine (graph already in pinned host memory).  At 16 GB the learned scheduler, the same
  code without its network (staging), the plain scheduler (the reference's
  max_choose_blocks) and the plain scheduler with the reference's zero-copy tail each run
  --warmup untimed and --steps timed walks, interleaved; SOWalker, the plain scheduler with
  the reference's upload format (beg and csr as stored, plain_raw; every other scheduler
  sends blocks in the transfer encoding) and the 8 GB cap run one timed walk per scheduler.
  value = steps/s of the learned scheduler at 16 GB; the others, H2D and zero-copy bytes
  are in config.caps.

Other workloads (--config): sopr_chung_lu (SOPR on a directed Chung-Lu graph, 41.6M V /
1.47B arcs, 4 walkers/vertex, term 0.15, fully HBM-resident), node2vec_rmat20 (the
reference's CPU-runnable case, single resident block), mixed_rmat27_ooc (Node2vec + SOSR
out of core, RMAT scale 27, >= 5.5B arcs) and node2vec_rmat24_ooc (reduced out-of-core case).

Prints one JSON line: walk steps/s plus the roofline of the step kernel (algorithmic bytes
per launch over its HIP-event time vs the 8 TB/s HBM peak) and a CPU baseline (the
reference's CPU step logic transcribed to C, on a bounded sample, all host cores).
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0

CONFIGS = {
    "sopr_chung_lu": dict(
        workload="SOPR, directed Chung-Lu 41.6M V / ~1.47B arcs (gamma 2.2, max out-degree ~3M, duplicate "
                 "targets and self-loops removed), 4 walkers/vertex, alpha 0.2, term 0.15, length cap 30, "
                 "fully HBM-resident",
        graph=dict(kind="chung_lu", nverts=41_600_000, nedges=1_470_000_000, gamma=2.2, max_degree=3_000_000),
        types=[dict(algo=1, length=30, alpha=0.2, term=0.15)], wpv=4),
    "node2vec_rmat26_ooc": dict(
        workload="Node2vec out of core, RMAT scale 26 (67.1M ids), edges drawn until >= 3.6B arcs remain after "
                 "symmetrizing and dedup, random id permutation, 256 MB blocks, HBM cap 16 GB (and 8 GB), p 0.5, q 2, "
                 "1 walker/vertex, length 40; learned vs plain block scheduler",
        graph=dict(kind="rmat", scale=26, nedges=1_856_000_000, min_arcs=3_600_000_000, undirected=True,
                   permute=True),
        types=[dict(algo=0, length=40, p=0.5, q=2.0)], wpv=1, out_of_core=True, block_mb=256, memory_gb=[16, 8]),
    "node2vec_rmat24_ooc": dict(
        workload="Node2vec out of core (reduced), RMAT scale 24 with 400M generated edges symmetrized, dedup, "
                 "random id permutation, 64 MB blocks, HBM cap 3 GB, p 0.5, q 2, 1 walker/vertex, length 40; "
                 "plain vs learned block scheduler",
        graph=dict(kind="rmat", scale=24, nedges=400_000_000, undirected=True, permute=True),
        types=[dict(algo=0, length=40, p=0.5, q=2.0)], wpv=1, out_of_core=True, block_mb=64, memory_gb=[3]),
    "mixed_rmat27_ooc": dict(
        workload="Mixed batch out of core: RMAT scale 27 (134.2M ids), directed edges drawn until >= 5.5B arcs "
                 "remain after dedup, ids in generation order (locality kept), 256 MB blocks, HBM cap 16 GB; "
                 "2 walkers/vertex: Node2vec p 1, q 0.25, length 80 and SOSR alpha 0.2, termination 0.2 (length "
                 "cap 80) co-scheduled in one kernel",
        graph=dict(kind="rmat", scale=27, nedges=5_630_000_000, min_arcs=5_500_000_000, undirected=False,
                   permute=False),
        types=[dict(algo=0, length=80, p=1.0, q=0.25), dict(algo=1, length=80, alpha=0.2, term=0.2)], wpv=2,
        out_of_core=True, block_mb=256, memory_gb=[16]),
    "node2vec_rmat20": dict(
        workload="Node2vec, RMAT scale 20 (2^20 V, 16M edges symmetrized, dedup, no self-loops), "
                 "p 2, q 0.5, 1 walker/vertex, length 80, single resident block",
        graph=dict(kind="rmat", scale=20, nedges=16 * 1024 * 1024, undirected=True, permute=False),
        types=[dict(algo=0, length=80, p=2.0, q=0.5)], wpv=1),
}


def search_mode(launch):
    """The step kernels' search-mode template argument as rocprof prints it: bit 0
    interpolated guesses, bit 1 window reads."""
    launch = launch or {}
    return "%du" % ((1 if launch.get("interp") else 0) | (2 if launch.get("window") else 0))


def algorithmic_bytes(st, nwalkers, visits):
    """Bytes the step kernel must move by the algorithm's definition:
    per walker: worklist id 4 + state read 12 (cur, prev, meta) + state write 12 + the two
    adjacency rows (beg[v], beg[v+1], beg[u], beg[u+1]) 32; per step: next row 16 (+ visit
    RMW 8); per trial: candidate 4; per probe: 4."""
    b = nwalkers * (4 + 12 + 12 + 32) + st["steps"] * (16 + (8 if visits else 0))
    b += st["trials"] * 4 + st["probes"] * 4
    return b


def sector_bytes(st, nwalkers, visits):
    """The same accesses priced at the 64-byte granularity of a random HBM touch: every
    row lookup, candidate and probe is one sector, a visit RMW two (read + write-back);
    walker state streams (worklist id, state read and written: 28 B per walker)."""
    touches = st["steps"] * (1 + (2 if visits else 0)) + st["trials"] + st["probes"] + nwalkers * 2
    return touches * 64 + nwalkers * 28


def build_graph(ctx, g, seed):
    """The configuration's graph in pinned host memory, and the edge count drawn.  RMAT
    draws `nedges` edges; when fewer than `min_arcs` arcs survive deduplication it draws
    again, scaled by the shortfall (the generator is deterministic in (seed, edge index),
    so a larger draw extends the smaller one)."""
    from gowalker_amd import engine as E
    if g["kind"] == "chung_lu":
        return E.gen_chung_lu(ctx, g["nverts"], g["nedges"], g["gamma"], g["max_degree"], seed, pinned=True) + \
            (g["nedges"],)
    drawn = g["nedges"]
    while True:
        beg, csr = E.gen_rmat(ctx, g["scale"], drawn, seed=seed, undirected=g["undirected"], permute=g["permute"],
                              pinned=True)
        if csr.n >= g.get("min_arcs", 0):
            return beg, csr, drawn
        print("[bench] %d edges drawn left %d arcs (< %d): drawing more" % (drawn, csr.n, g["min_arcs"]),
              file=sys.stderr, flush=True)
        drawn = int(drawn * g["min_arcs"] / csr.n * 1.002) + 1
        beg.close()
        csr.close()


def cpu_baseline(beg, csr, cfg, seconds, threads):
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import numpy as np

    import _oracle as O
    t = cfg["types"][0]
    n = len(beg) - 1
    rng = np.random.default_rng(7)
    sample = 4_000_000
    src = rng.integers(0, n, sample).astype(np.uint32)
    algo = O.N2V if t["algo"] == 0 else O.SOPR
    def timed(literal, nthreads, secs, nwalk):
        t0 = time.perf_counter()
        r = O.ref_run(beg, csr, algo, t["length"], p=t.get("p", 1.0), q=t.get("q", 1.0), alpha=t.get("alpha", 0.2),
                      term=t.get("term", 0.0), literal=literal, nwalkers=nwalk, sources=src[:nwalk], max_seconds=secs,
                      nthreads=nthreads)
        return r, time.perf_counter() - t0

    r, dt = timed(False, threads, seconds, sample)
    r1, dt1 = timed(False, 1, seconds / 4, sample // 8)
    rl, dtl = timed(True, threads, seconds / 4, sample)
    # this build's own algorithm on the host cores: the oracle's Philox replay (same walks as
    # the GPU), on a walker count sized from a short trial run to about seconds / 4
    if t["algo"] == 0:
        tc = O.type_cfg(O.N2V, t["length"], p=t.get("p", 1.0), q=t.get("q", 1.0), term=t.get("term", 0.0))
    else:
        tc = O.type_cfg(O.SOPR, t["length"], alpha=t.get("alpha", 0.2), term=t.get("term", 0.0))
    wc = O.walk_cfg(42, tc)
    nw = 65536
    t0 = time.perf_counter()
    O.run(beg, csr, wc, nw, sources=src[:nw], state=False, nthreads=threads)
    nw = int(min(sample, max(nw, nw * (seconds / 4) / max(time.perf_counter() - t0, 1e-3))))
    t0 = time.perf_counter()
    rp = O.run(beg, csr, wc, nw, sources=src[:nw], state=False, nthreads=threads)
    dtp = time.perf_counter() - t0
    return dict(value=r["moves"] / dt, unit="steps/s", cores=threads, kind="port",
                sample="%d of %d random-source walkers run to completion in %.1f s (%d moves); C transcription of "
                       "the reference CPU step (engine/engine.hpp:158-404, fixed-mode RandNum), OpenMP"
                       % (r["done"], sample, dt, r["moves"]),
                value_1core=r1["moves"] / dt1,
                value_literal_randnum=rl["moves"] / dtl,
                value_philox_replay=rp["totals"]["steps"] / dtp,
                note="value_1core: same code on one thread (%d walkers); value_literal_randnum: the reference's "
                     "type-punned inlineRandPair (util/util.hpp:291-301), nearly first-order sampling; "
                     "value_philox_replay: this build's walk (Philox draws, integer thresholds, the GPU's samplers) "
                     "replayed on the same cores (%d walkers)" % (r1["done"], nw))


LDS_BLOCKS = 128          # block tables up to this size are staged in LDS by the step kernel
PCIE_PEAK_GBS = 64.0   # PCIe Gen5 x16, one direction

# name -> (scheduler, zero-copy threshold rule, gwh flags); the zero-copy threshold of
# plain_zero is the reference's `zero` option (blocksize/128/24 walkers, apps/Node2vec.cu:100-104)
SCHEDULERS = {
    "learned": ("SCHED_LEARNED", False, 0),         # the MLP scheduler (network carried across walks)
    "staging": ("SCHED_STAGING", False, 0),         # the same staging/prefetch/tail code, no network
    "plain": ("SCHED_PLAIN", False, 0),             # the reference's max_choose_blocks, no zero-copy tail
    "plain_zero": ("SCHED_PLAIN", True, 0),         # ... with the reference's zero-copy tail
    "plain_raw": ("SCHED_PLAIN", False, 0),         # the reference's scheduler and upload format
                                                    # (beg and csr as stored, gpu/memory.hpp:124-139)
    "sowalker": ("SCHED_SOWALKER", False, 0),       # the reference's simulated-annealing scheduler
    "learned_model_swap": ("SCHED_LEARNED", False, 4),        # ablation: the network picks swap loads
    "learned_model_swap_myopic": ("SCHED_LEARNED", False, 6),  # ... trained on next-run labels
    "learned_myopic": ("SCHED_LEARNED", False, 2),            # ablation: next-run labels, plain swaps
    "learned_value_victim": ("SCHED_LEARNED", False, 12),     # ... and victims by value
    "learned_no_swap_prefetch": ("SCHED_LEARNED", False, 1),  # ablation: no staging slot
    "staging_no_sticky": ("SCHED_STAGING", False, 16),        # ablation: no preference for the staged block
    "learned_two_lanes": ("SCHED_LEARNED", False, 128),       # ablation: packed uploads on two copy lanes
}
HEADLINE = ["learned", "staging", "plain", "plain_zero"]   # timed --warmup + --steps walks each
RAW = {"plain_raw"}                                        # uploads in the reference's format
ONCE = ["sowalker", "plain_raw"]                           # one timed walk at the headline cap
EXTRA = ["learned", "staging", "learned_model_swap", "plain_zero", "plain", "plain_raw", "sowalker"]   # one timed walk
                                                                            # each at other caps


def ooc_walk(E, args, cfg, beg, csr, nwalkers, cap_gb, name, models, timed, packed=None, hashes=False):
    """One full walk.  Every scheduler carries a model across its walks: the learned one
    its network, all of them the launch tuner, which tries a candidate launch per untimed
    walk and runs the fastest tried in timed walks.  Warm-up and timed walks keep no
    per-walker output (the reference's walks have none either); a check walk (hashes=True,
    untimed, after the timed ones) keeps the per-walker path hash, 8 B per walker counted
    against the cap."""
    sched, zero, flags = SCHEDULERS[name]
    if timed or hashes:
        flags |= E.TUNE_FREEZE
    block_bytes = cfg["block_mb"] << 20
    launch = None
    if args.variant is not None or args.waves_per_cu or args.window or args.block:
        # a fixed launch: with --block the search order is the flags given, else the engine's
        # probe model picks it
        launch = dict(block=args.block, variant=2 if args.variant is None else args.variant,
                      waves_per_cu=args.waves_per_cu, window=args.window, interp=args.interp)
    rep, h, _ = E.run_out_of_core(beg, csr, cfg["types"], args.seed, nwalkers, wpv=cfg["wpv"],
                                  memory_bytes=int(cap_gb * (1 << 30)), block_bytes=block_bytes,
                                  zero_threshold=block_bytes // 128 // 24 if zero else 0,
                                  scheduler=getattr(E, sched), hashes=hashes, flags=flags, launch=launch,
                                  prefetch_depth=args.prefetch_depth, verbose=args.verbose,
                                  model=models.setdefault(name, E.Model()),
                                  packed=None if name in RAW else packed)
    return rep, h


def summarize(reps, zt):
    r0 = reps[0]
    wall = sum(r["wall_s"] for r in reps)
    zc_gb = r0["zc_touches"] * 64 / 1e9
    return dict(reps=len(reps), steps_per_s=sum(r["steps"] for r in reps) / wall, ms_per_run=wall * 1e3 / len(reps),
                ms_min=min(r["wall_s"] for r in reps) * 1e3, ms_max=max(r["wall_s"] for r in reps) * 1e3,
                h2d_gb=r0["h2d_bytes"] / 1e9, h2d_raw_gb=r0["h2d_raw_bytes"] / 1e9, zero_copy_gb_est=zc_gb,
                pcie_gb_total=r0["h2d_bytes"] / 1e9 + zc_gb,
                h2d_gbs=(r0["h2d_bytes"] / 1e9) / (r0["h2d_ms"] / 1e3) if r0["h2d_ms"] else None,
                iterations=r0["iterations"], kernel_launches=r0["kernel_launches"], uploads=r0["uploads"],
                staged=r0["staged"],
                zero_copy_steps=r0["zero_copy_steps"], kernel_ms=r0["kernel_ms"], copy_ms=r0["h2d_ms"],
                sched_ms=r0["sched_ms"], learned_decisions=r0["learned_decisions"], fill_model=r0["fill_model"],
                swap_model=r0["swap_model"], swap_differs=r0["swap_differs"], tail_model=r0["tail_model"],
                tail_threshold=r0["tail_threshold"], model_samples=r0["model_samples"], fallbacks=r0["fallbacks"],
                nslots=r0["nslots"], nblocks=r0["nblocks"], steps=r0["steps"], trials=r0["trials"],
                probes=r0["probes"], zero_threshold=zt, launch=r0["tuned"], copy_lanes=r0["copy_lanes"],
                decoded_in_place=r0["inplace_blocks"], decoded_ring=r0["ring_blocks"],
                decoded_copy_stream=r0["inline_blocks"])


def log_walk(cap_gb, name, tag, rep):
    print("[bench] cap %g GB %s %s: %.3f s, %d steps, %.2f GB H2D (%.2f GB raw; %d uploads + %d staged), %.3f GB zero-copy, "
          "%d iterations (%d launches), kernel %.0f ms, copy %.0f ms, sched %.0f ms, %d slots / %d blocks, model: fill %d swap %d "
          "(%d differ) tail %d/%d samples %d, sticky %d, launch %s, copy lanes %d, decoded in place/ring/copy stream "
          "%d/%d/%d"
          % (cap_gb, name, tag, rep["wall_s"], rep["steps"], rep["h2d_bytes"] / 1e9, rep["h2d_raw_bytes"] / 1e9,
             rep["uploads"], rep["staged"],
             rep["zc_touches"] * 64 / 1e9, rep["iterations"], rep["kernel_launches"], rep["kernel_ms"], rep["h2d_ms"],
             rep["sched_ms"],
             rep["nslots"], rep["nblocks"], rep["fill_model"], rep["swap_model"], rep["swap_differs"],
             rep["tail_model"], rep["tail_threshold"], rep["model_samples"], rep["sticky_swaps"],
             "%(block)d/%(waves_per_cu)d/v%(variant)d" % rep["tuned"] if rep["tuned"] else "-", rep["copy_lanes"],
             rep["inplace_blocks"], rep["ring_blocks"], rep["inline_blocks"]),
          file=sys.stderr, flush=True)


def run_ooc(args, cfg, ctx, beg, csr, nverts, narcs, gen_s, t_start):
    """Out-of-core configuration: the C++ engine under each scheduler.  One timed step = one
    full walk of every walker.  At the headline cap the learned, staging, plain and
    plain_zero schedulers each run --warmup untimed and --steps timed walks, interleaved
    walk by walk; SOWalker and the other caps run one timed walk each (the learned one
    after --extra-warmup untimed walks).  A wall-clock guard (--time-budget) stops adding
    headline walks when the budget would be overrun, and the line reports the walks done."""
    import numpy as np

    from gowalker_amd import engine as E
    ctx.close()
    caps = args.caps or cfg["memory_gb"]
    block_bytes = cfg["block_mb"] << 20
    zt = block_bytes // 128 // 24
    nwalkers = nverts * cfg["wpv"]
    models = {}      # one network per learned variant, carried across its walks
    ref_hash = None
    packed, pack = None, None
    if not args.no_pack:
        # the blocks in the transfer encoding, built once on the host cores (not timed, like
        # the graph's preprocessing into blocks)
        t0 = time.perf_counter()
        packed = E.PackedGraph(beg, csr, block_bytes)
        pack = dict(packed.info(), pack_s=time.perf_counter() - t0)
        pack["ratio"] = pack["packed_bytes"] / pack["raw_bytes"]
        print("[bench] packed %d of %d blocks: %.2f GB of csr sent as %.2f GB (%.3f) in %.1f s"
              % (pack["blocks_packed"], pack["nblocks"], pack["raw_bytes"] / 1e9, pack["packed_bytes"] / 1e9,
                 pack["ratio"], pack["pack_s"]), file=sys.stderr, flush=True)
    identical = True

    def check_hash(h):
        nonlocal ref_hash, identical
        if ref_hash is None:
            ref_hash = h
        elif not np.array_equal(h, ref_hash):
            identical = False

    def check_walks(cap_gb, names):
        """one untimed walk with per-walker path hashes per scheduler: walks must be
        bit-identical across schedulers and caps"""
        for name in names:
            rep, h = ooc_walk(E, args, cfg, beg, csr, nwalkers, cap_gb, name, models, False, packed, hashes=True)
            check_hash(h)
            log_walk(cap_gb, name, "check", rep)

    head = [n for n in (args.schedulers or HEADLINE)]
    cap0 = caps[0]
    reps = {n: [] for n in head}
    done_rounds = 0
    round_s = []
    # later work: the CPU baseline, check walks at the headline cap, and each further cap's
    # walks and check walks
    reserve = (0 if args.no_cpu_baseline else args.cpu_seconds * 2.2 + 5) + 10 + (80 * (len(caps) - 1))
    for i in range(args.warmup + args.steps):
        elapsed = time.perf_counter() - t_start
        est = max(round_s) if round_s else 0.0
        if i >= args.warmup and done_rounds >= 1 and elapsed + est + reserve > args.time_budget:
            print("[bench] time budget: %.0f s used, stopping after %d timed walks per scheduler"
                  % (elapsed, done_rounds), file=sys.stderr, flush=True)
            break
        t0 = time.perf_counter()
        for name in head:
            rep, h = ooc_walk(E, args, cfg, beg, csr, nwalkers, cap0, name, models, i >= args.warmup, packed)
            log_walk(cap0, name, "warmup %d" % i if i < args.warmup else "run %d" % (i - args.warmup), rep)
            if i >= args.warmup:
                reps[name].append(rep)
        round_s.append(time.perf_counter() - t0)
        if i >= args.warmup:
            done_rounds += 1
    per_cap = []
    res0 = {n: summarize(reps[n], zt if SCHEDULERS[n][1] else 0) for n in head if reps[n]}
    tuning = {}
    for n in head:
        rows, ch = models[n].tuning()
        tuning[n] = dict(chosen=ch, candidates=[dict(r, kernel_ms_per_walk=r["ms_per_gstep"] * reps[n][0]["steps"] / 1e9
                                                     if reps[n] else None) for r in rows])
    if not args.schedulers:
        for name in ONCE:
            rep, h = ooc_walk(E, args, cfg, beg, csr, nwalkers, cap0, name, models, True, packed)
            log_walk(cap0, name, "run 0", rep)
            res0[name] = summarize([rep], 0)
    check_walks(cap0, list(res0))
    per_cap.append({"memory_gb": cap0, "schedulers": res0})
    for cap_gb in caps[1:]:
        res = {}
        for name in (args.schedulers or EXTRA):
            if name.startswith("learned"):
                for w in range(args.extra_warmup):
                    rep, h = ooc_walk(E, args, cfg, beg, csr, nwalkers, cap_gb, name, models, False, packed)
                    log_walk(cap_gb, name, "warmup %d" % w, rep)
            rep, h = ooc_walk(E, args, cfg, beg, csr, nwalkers, cap_gb, name, models, True, packed)
            log_walk(cap_gb, name, "run 0", rep)
            res[name] = summarize([rep], zt if SCHEDULERS[name][1] else 0)
        check_walks(cap_gb, list(res))
        per_cap.append({"memory_gb": cap_gb, "schedulers": res})
    for entry in per_cap:
        res = entry["schedulers"]
        for other in ("staging", "plain", "plain_zero", "plain_raw", "sowalker"):
            if "learned" in res and other in res:
                entry["learned_vs_%s_steps_per_s" % other] = res["learned"]["steps_per_s"] / res[other]["steps_per_s"]
                entry["learned_vs_%s_pcie_bytes" % other] = res["learned"]["pcie_gb_total"] / res[other]["pcie_gb_total"]
        if "staging" in res and "plain" in res:
            entry["staging_vs_plain_steps_per_s"] = res["staging"]["steps_per_s"] / res["plain"]["steps_per_s"]
    lead = "learned" if "learned" in res0 else head[0]
    headrec = res0[lead]
    r0 = reps[lead][0]
    launches = max(1, r0["kernel_launches"])   # iterations plus worklist-overflow rounds
    abytes = algorithmic_bytes(r0, nwalkers, False) / launches      # per launch, averaged
    achieved = abytes / (r0["kernel_ms"] / launches * 1e-3) / 1e9
    traffic, traffic_note = None, None
    if os.path.exists(args.traffic_json):
        try:
            tj = json.load(open(args.traffic_json)).get(args.config)
        except Exception:
            tj = None
        if tj and tj.get("bytes_per_step"):
            traffic = tj["bytes_per_step"] * r0["steps"] / launches
            traffic_note = tj.get("source")
    mi = {n: m.info() for n, m in models.items() if n.startswith("learned")}
    if packed is not None:
        packed.close()
    for m in models.values():
        m.close()
    line = {
        "metric": "walk steps/s on 1 MI355X + H2D GB, learned vs plain block scheduler",
        "value": headrec["steps_per_s"],
        "unit": "steps/s",
        "n_gpus": 1, "steps": done_rounds, "warmup": args.warmup,
        "ms_per_step": headrec["ms_per_run"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
        "data": "synthetic (generated on the GPU, seed %d)" % args.seed,
        "config": {"workload": cfg["workload"], "nverts": nverts, "narcs": narcs,
                   "non_isolated": int(np.count_nonzero(beg[1:] != beg[:-1])), "walkers": nwalkers,
                   "edges_drawn": cfg.get("edges_drawn"), "caps": per_cap, "walks_identical": identical,
                   "walk_check": "timed and warm-up walks keep no per-walker output; after them one untimed walk per "
                                 "scheduler and cap keeps the per-walker path hash (8 B per walker, counted against the "
                                 "cap), and walks_identical compares those hashes",
                   "model": mi, "launch_tuning": tuning, "packed_uploads": pack, "graph_gen_s": gen_s, "pcie_peak_gbs": PCIE_PEAK_GBS, "parallelism": "1 GPU",
                   "steps_requested": args.steps, "launch": r0.get("tuned")},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_over_algorithmic": traffic / abytes if traffic else None,
                     # the lane kernel's last argument: walkers whose previous block was evicted
                     # (runs after the first eviction of a walk; none where every block fits)
                     "kernel": ("walk_lane_kernel<false, %s, %s, %s, " + ("true>" if r0["nslots"] < r0["nblocks"]
                                                                            else "false>")
                                if (r0.get("tuned") or {}).get("variant") == 2 else "walk_kernel<false, %s, %s, %s>")
                               % (search_mode(r0.get("tuned")), "true" if r0["nblocks"] <= LDS_BLOCKS else "false",
                                  "true" if len(cfg["types"]) == 2 else "false"),
                     "kernel_ms": r0["kernel_ms"] / launches, "algorithmic_bytes": abytes,
                     "sector_bytes": sector_bytes(r0, nwalkers, False) / launches, "launches": launches,
                     "per_step": {"steps_per_launch": r0["steps"] / launches,
                                  "algorithmic_bytes": algorithmic_bytes(r0, nwalkers, False) / r0["steps"],
                                  "traffic_bytes": (traffic * launches / r0["steps"]) if traffic else None},
                     "traffic_source": traffic_note,
                     "note": "per launch, averaged over the step-kernel launches of the first timed %s walk at %g GB; "
                             "traffic = HBM bytes per step from rocprofv3 FETCH_SIZE x 2 + WRITE_SIZE (profiles/) "
                             "times this walk's steps per launch" % (lead, cap0)},
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        threads = len(os.sched_getaffinity(0))
        line["cpu_baseline"] = cpu_baseline(beg, csr, cfg, args.cpu_seconds, threads)
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="node2vec_rmat26_ooc", choices=sorted(CONFIGS))
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="CPU baseline: seconds of the main all-core sample (the three side samples take a "
                         "quarter each)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--waves-per-cu", type=int, default=0)
    ap.add_argument("--variant", type=int, default=None,
                    help="step kernel: 0 per-step lane loop, 1 flattened state machine, 2 lean per-load lanes "
                         "(default: the autotuner's pick for resident workloads, 2 out of core)")
    ap.add_argument("--prefetch-depth", type=int, default=0,
                    help="out-of-core learned scheduler: staged block copies in flight (0: the library's 8)")
    ap.add_argument("--verbose", action="store_true", help="out-of-core: the engine logs every iteration to stderr")
    ap.add_argument("--interp", action="store_true",
                    help="resident workloads with a fixed shape: membership by interpolation search")
    ap.add_argument("--window", action="store_true",
                    help="membership probes read aligned 64-byte windows (resident: with a fixed shape)")
    ap.add_argument("--traffic-json", default=os.path.join(REPO, "profiles", "traffic.json"))
    ap.add_argument("--caps", type=lambda s: [float(x) for x in s.split(",")], default=None,
                    help="out-of-core: comma list of HBM caps in GB (the first is the headline cap)")
    ap.add_argument("--schedulers", type=lambda s: s.split(","), default=None,
                    help="out-of-core: comma list of schedulers (%s), run at every cap; default: %s timed "
                         "at the headline cap, SOWalker once, %s once at other caps"
                         % (",".join(SCHEDULERS), ",".join(HEADLINE), ",".join(EXTRA)))
    ap.add_argument("--no-pack", action="store_true",
                    help="out-of-core: upload blocks as stored (4 bytes per arc) instead of in the transfer encoding")
    ap.add_argument("--block-mb", type=int, default=0,
                    help="out-of-core: graph block size in MB (default: the configuration's, 256)")
    ap.add_argument("--extra-warmup", type=int, default=2,
                    help="out-of-core: untimed learned walks before the timed one at caps after the first")
    ap.add_argument("--time-budget", type=float, default=400.0,
                    help="out-of-core: seconds after which no further headline walk starts once one timed round is "
                         "done (the line still prints)")
    ap.add_argument("--autotune", type=int, default=4_000_000,
                    help="resident workloads: time every launch shape and search order on this many walkers "
                         "(untimed, state restored) and run with the fastest; 0 = fixed shape")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one MI355X (--gpus 1); the engine runs one process per GPU, and sharding by "
              "walker id (resident graphs) or by block (out of core) is not run by this benchmark.",
              file=sys.stderr)
        sys.exit(2)

    from gowalker_amd import engine as E
    t_start = time.perf_counter()
    cfg = dict(CONFIGS[args.config])
    if args.block_mb:
        cfg["block_mb"] = args.block_mb
    ctx = E.Context(0)
    t0 = time.perf_counter()
    beg, csr, cfg["edges_drawn"] = build_graph(ctx, cfg["graph"], args.seed)
    gen_s = time.perf_counter() - t0
    nverts = beg.n - 1
    narcs = csr.n
    print("[bench] graph %s: %d vertices, %d arcs (%d edges drawn), generated in %.2f s"
          % (args.config, nverts, narcs, cfg["edges_drawn"], gen_s), file=sys.stderr, flush=True)
    if cfg.get("out_of_core"):
        run_ooc(args, cfg, ctx, beg.a, csr.a, nverts, narcs, gen_s, t_start)
        beg.close()
        csr.close()
        return
    block_bytes = 1 << 34   # one block holds the whole graph (fully resident)
    blocks = E.Blocks(beg.a, csr.a, block_bytes)
    ctx.set_blocks(blocks)
    t0 = time.perf_counter()
    ctx.load_all()
    load_s = time.perf_counter() - t0
    h2d_graph = ctx.counters()["h2d_bytes"]
    ctx.set_walk(cfg["types"], seed=args.seed)
    nwalkers = nverts * cfg["wpv"]
    visits = cfg["types"][0]["algo"] == 1
    outputs = E.OUT_VISITS if visits else E.OUT_HASH
    if args.variant is None:
        args.variant = 0
    launch = dict(block=args.block, waves_per_cu=args.waves_per_cu, variant=args.variant, interp=args.interp,
                  window=args.window)

    def one_step():
        ctx.init_walkers(nwalkers, cfg["wpv"], outputs=outputs)
        st = ctx.run(**launch)
        assert st["walkers_left"] == 0, st
        return st

    tuned = None
    if args.autotune:
        ctx.init_walkers(nwalkers, cfg["wpv"], outputs=outputs)
        tuned, tms = ctx.autotune(min(args.autotune, nwalkers))
        launch = dict(block=0)
        args.variant = tuned["variant"]
        launch = dict(block=0)
        print("[bench] autotune: %s (%.2f ms on %d walkers)" % (tuned, tms, min(args.autotune, nwalkers)),
              file=sys.stderr, flush=True)

    for i in range(args.warmup):
        st = one_step()
        print("[bench] warmup %d: %d steps, kernel %.1f ms" % (i, st["steps"], st["kernel_ms"]), file=sys.stderr,
              flush=True)
    stats = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        stats.append(one_step())
    dt = time.perf_counter() - t0
    steps = sum(s["steps"] for s in stats)
    kms = [s["kernel_ms"] for s in stats]
    st0 = stats[0]
    for s in stats[1:]:
        assert s["steps"] == st0["steps"] and s["trials"] == st0["trials"], "walks differ between steps"
    avg_kms = sum(kms) / len(kms)
    abytes = algorithmic_bytes(st0, nwalkers, visits)
    achieved = abytes / (avg_kms * 1e-3) / 1e9
    traffic, traffic_note = None, None
    if os.path.exists(args.traffic_json):
        try:
            tj = json.load(open(args.traffic_json)).get(args.config)
        except Exception:
            tj = None
        if tj and tj.get("bytes_per_step"):
            traffic = tj["bytes_per_step"] * st0["steps"]       # one launch per pass
            traffic_note = tj.get("source")
    line = {
        "metric": "walk steps/s on 1 MI355X",
        "value": steps / dt,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (generated on the GPU, seed %d)" % args.seed,
        "config": {"workload": cfg["workload"], "nverts": nverts, "narcs": narcs, "walkers": nwalkers,
                   "walk_steps_per_pass": st0["steps"], "trials_per_pass": st0["trials"],
                   "probes_per_pass": st0["probes"], "mixture_steps_per_pass": st0["mix_steps"],
                   "h2d_gb": 0.0, "graph_h2d_gb_untimed": h2d_graph / 1e9, "graph_load_s": load_s,
                   "parallelism": "1 GPU", "launch": tuned or launch},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_over_algorithmic": traffic / abytes if traffic else None,
                     "per_step": {"algorithmic_bytes": abytes / st0["steps"],
                                  "traffic_bytes": traffic / st0["steps"] if traffic else None},
                     "traffic_source": traffic_note,
                     "kernel": "walk_flat_kernel<true>" if args.variant == 1 else
                               ("walk_lane_kernel<true, %s, false, %s, false>" if args.variant == 2
                                else "walk_kernel<true, %s, false, %s>")
                               % (search_mode(tuned or launch), "true" if len(cfg["types"]) == 2 else "false"),
                     "kernel_ms": avg_kms, "algorithmic_bytes": abytes,
                     "sector_bytes": sector_bytes(st0, nwalkers, visits)},
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        threads = len(os.sched_getaffinity(0))
        line["cpu_baseline"] = cpu_baseline(beg.a, csr.a, cfg, args.cpu_seconds, threads)
    print(json.dumps(line), flush=True)
    ctx.close()
    beg.close()
    csr.close()


if __name__ == "__main__":
    main()
