#!/usr/bin/env python3
"""Label-rate benchmark of the MI355X PoSE-DB labeler.

Default workload (the headline): batched persistent labeling at
chunk_blocks = 2^17.  One "step" is one pass of the hot path over one batch:
a single persistent launch, replayed from a captured hipGraph, that labels
`--super` super-chunks (keyed BLAKE3, exact integer arithmetic) into a device
buffer already resident in HBM.  Session key: SplitMix64 key of PRNG seed 2;
start: an aligned absolute position drawn from the same stream.  W warmup
steps, then K timed steps between device synchronisations.

The same JSON line also carries the plain form (one super-chunk graph per
launch) and the 64-super-chunk batch at 2^17 (latency-bound), so batched vs
plain vs CPU are measured in one run.  Every run re-checks sampled
super-chunks of the GPU output against the CPU labeler and times the
reference CPU labeler (compiled from source under oracle/_ref) on a bounded
sample of the same workload on the host.

Other workloads (`--workload`): plain (256 super-chunks at 2^14, one per
launch, full byte compare), uneven (48 sessions of mixed chunk_blocks in one
hipGraph launch), serve (Poisson request stream through the label server,
p50/p99 per-super-chunk latency), verify (challenge answers per second, each
labeling a whole super-chunk graph, against the reference's CPU challenge).
"""
import argparse
import json
import math
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "pose-memory-wiping_amd"))
sys.path.insert(0, os.path.join(REPO, "oracle"))

from workloads import SplitMix, poisson_requests, uneven_sessions  # noqa: E402

# int32 VALU ops executed per scaffold node by the labeling kernel: two BLAKE3
# compressions (block 1 and block 2), 7 rounds x 8 G x 12 instructions + 8
# output XORs each.  Block 0 is part of the precomputed topology.
OPS_PER_NODE = 2 * (7 * 8 * 12 + 8)
MIB = 1 << 20
# HBM-side bytes per scaffold node of the batched launch at chunk_blocks=2^17,
# 64 lanes, list schedule, keyed by (chunk_blocks, lanes, schedule, waves per
# SIMD): FETCH_SIZE and WRITE_SIZE passes scaled by a copy kernel of known size
# (profiles/k_label_pipe_cb131072_8192sc_counters.md at 4 waves per SIMD,
# profiles/k_label_cb131072_8192sc_counters.md at 8).
TRAFFIC_BYTES_PER_NODE = {(1 << 17, 64, 1, 4): 90.5, (1 << 17, 64, 1, 8): 106.4}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def valu_peak(props):
    cus = props.multi_processor_count
    # MI355X: 64 int32 lanes per CU per clock, 2.4 GHz peak engine clock.
    return cus * 64 * 2.4e9, cus


def cpu_expected(seed, cb, js, rbo, threads):
    """Expected super-chunks js: the reference build when present (fast), else the oracle."""
    from pyoracle import Oracle, RefLib, ref_available
    js = list(js)
    if ref_available(threads) and js == list(range(js[0], js[0] + len(js))):
        ref = RefLib(threads)
        try:
            return ref.label_region(seed, cb, len(js), rbo + js[0] * cb)
        finally:
            ref.close()
    return Oracle().label_superchunks(seed, cb, js, rbo, threads)


def cpu_expected_many(items, threads):
    """Expected labels of many single super-chunks (session seed, chunk_blocks,
    absolute block of the super-chunk), labeled concurrently on `threads` host
    threads by the reference build (the oracle when it is absent)."""
    import threading
    from concurrent.futures import ThreadPoolExecutor
    from pyoracle import Oracle, RefLib, ref_available
    local, made, lock = threading.local(), [], threading.Lock()
    use_ref = ref_available(4)

    def one(item):
        seed, cb, block = item
        if use_ref:
            if not hasattr(local, "ref"):
                local.ref = RefLib(4)
                with lock:
                    made.append(local.ref)
            return local.ref.label_region(seed, cb, 1, block)
        return Oracle().label_superchunks(seed, cb, [0], block, 1)

    with ThreadPoolExecutor(max_workers=max(1, threads)) as ex:
        out = list(ex.map(one, list(items)))
    for r in made:
        r.close()
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def cpu_baseline(cb, key, rbo, n_sample, threads):
    from pyoracle import Oracle, RefLib, ref_available
    if ref_available(threads):
        ref = RefLib(threads)
        t0 = time.perf_counter()
        ref.label_region(key, cb, n_sample, rbo)
        dt = time.perf_counter() - t0
        ref.close()
        kind = "reference"
    else:
        t0 = time.perf_counter()
        Oracle().label_region(key, cb, n_sample, rbo, threads)
        dt = time.perf_counter() - t0
        kind = "port"
    return dt, kind


def reference_topology_build(cb):
    """The reference's host-side topology work for one chunk_blocks (its edge
    enumeration and slot map), timed on this host; None without the reference
    build."""
    from pyoracle import RefLib, ref_available
    if not ref_available(4):
        return None
    ref = RefLib(4)
    e_s, s_s, count = ref.topology_build_seconds(cb)
    ref.close()
    return {"edges_s": round(e_s, 2), "slot_map_s": round(s_s, 2), "nodes": count,
            "sample": f"pose_graph_edges + pose_graph_slot_map at chunk_blocks={cb}, the host calls of the "
                      f"reference's topology build (label_gpu_common.cu:91, :173), on one thread as there"}


def check_sample(out, n_super, cb, key, rbo, js, threads):
    super_bytes = cb * 32
    host = out.view(n_super, super_bytes)
    for j in js:
        exp = cpu_expected(key, cb, [j], rbo, threads)
        if host[j].cpu().numpy().tobytes() != exp:
            print(f"PARITY FAILURE: super-chunk {j} differs from the CPU labeler", file=sys.stderr)
            sys.exit(1)
    log(f"parity: super-chunks {js} byte-identical to the CPU labeler")


def launch_shape(pg, sessions, slots):
    """Form the automatic plan chooses for these sessions (read from a plan
    built outside any timed region): lanes per node and workgroup jobs per
    super-chunk (2 when each scaffold copy gets its own workgroup)."""
    plan = pg.Plan(sessions, slots, 0, graph=False)
    g = plan.groups()[0]
    plan.close()
    return {"lanes_per_node": g["lanes_per_node"],
            "workgroups_per_superchunk": g["jobs"] // max(1, g["super_chunks"])}


def workload_batched(args, torch, pg, dev, peak, cus):
    sm = SplitMix(2)
    key = sm.key()
    cb = args.cb
    rbo = sm.aligned_start(cb)
    tuned = None
    if args.autotune:
        n_target = args.super or cus * 32 // (args.lanes // 64)
        t0 = time.perf_counter()
        tuned, trials = pg.autotune(cb, n_target)
        log(f"autotune ({time.perf_counter() - t0:.1f} s): " +
            "; ".join(f"lanes {t['lanes']} wps {t['waves_per_simd']}: {t['batch_mib_per_s']:.0f} MiB/s"
                      for t in trials))
        args.lanes, args.mode, args.waves_per_simd = tuned["lanes"], tuned["mode"], tuned["waves_per_simd"]
        args.lanes_per_node = tuned["lanes_per_node"]
    t0 = time.perf_counter()
    topo = pg.Topology(cb, args.lanes, args.mode)
    build_s = time.perf_counter() - t0
    info = topo.info
    log(f"topology cb={cb} lanes={info.lanes} steps={info.steps} levels={info.levels} "
        f"pool={info.pool_slots} build={build_s:.2f}s")

    # Eight waves' worth of super-chunks per SIMD; at the default four waves
    # per SIMD each workgroup slot labels two of them in turn.
    n_super = args.super or cus * 32 // (args.lanes // 64)
    super_bytes = cb * 32
    out = torch.empty(n_super * super_bytes, dtype=torch.uint8, device=dev)
    plan = pg.Plan([(topo, key, rbo, n_super, out.data_ptr())], args.slots, args.waves_per_simd,
                   graph=True, lanes_per_node=args.lanes_per_node)
    pi = plan.info
    log(f"plan: {pi.super_chunks} super-chunks, {pi.slots} slots, {pi.lanes_per_node} lane(s) per node, "
        f"graph={pi.uses_graph}, scratch {pi.scratch_bytes / 2**30:.1f} GiB")
    for w in range(args.warmup):
        ms = plan.run()
        log(f"warmup {w}: kernel {ms:.1f} ms")
    torch.cuda.synchronize()
    kms = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        kms.append(plan.run())
        log(f"step {k}: kernel {kms[-1]:.1f} ms")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    mibs = n_super * super_bytes * args.steps / elapsed / MIB
    kernel_s = sum(kms) / len(kms) / 1e3
    nodes = info.nodes * n_super
    achieved = nodes * OPS_PER_NODE / kernel_s
    plan.close()

    js = sorted({0, n_super - 1, n_super // 2})
    check_sample(out, n_super, cb, key, rbo, js, args.cpu_threads)

    extras = {}
    if not args.no_extras:
        # Plain form: one super-chunk graph per launch, sequential launches.
        # One untimed launch first: the first launch that labels the scaffold
        # copies on their own uploads their schedules (once per topology).
        pg.label_batch([(topo, key, rbo, 1, out.data_ptr())], 1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for j in range(args.plain_launches):
            pg.label_batch([(topo, key, rbo + j * cb, 1, out.data_ptr() + j * super_bytes)], 1)
        torch.cuda.synchronize()
        plain_s = time.perf_counter() - t0
        plain_mibs = args.plain_launches * super_bytes / plain_s / MIB
        log(f"plain: {args.plain_launches} launches of one super-chunk in {plain_s:.2f} s")
        # 64 super-chunks in one batched launch (latency-bound at this depth).
        n64 = 64
        t0 = time.perf_counter()
        ms64 = pg.label_batch([(topo, key, rbo, n64, out.data_ptr())], n64)
        torch.cuda.synchronize()
        b64_s = time.perf_counter() - t0
        log(f"batch of 64: kernel {ms64:.1f} ms, wall {b64_s:.2f} s")
        streamed = streamed_batch_through_server(args, pg, cus)
        plain_shape = launch_shape(pg, [(topo, key, rbo, 1, out.data_ptr())], 1)
        b64_shape = launch_shape(pg, [(topo, key, rbo, n64, out.data_ptr())], n64)
        del out
        torch.cuda.empty_cache()
        # Uneven sessions in one hipGraph (autotuned per chunk_blocks) and the
        # serving loop at three arrival rates, each with its own parity sample.
        u_mibs, u_ms, _, u_nodes, u_detail = measure_uneven(args, torch, pg, dev, args.uneven_steps, 1)
        s_rates = [float(r) for r in args.extra_rates.split(",")]
        s_results, s_checked = measure_serving(args, pg, s_rates, args.extra_serve_seconds)
        extras = {"plain_one_superchunk_per_launch": dict(
                      {"value": round(plain_mibs, 3), "unit": "MiB/s", "launches": args.plain_launches,
                       "ms_per_launch": round(plain_s / args.plain_launches * 1e3, 1)}, **plain_shape),
                  "batched_vs_plain": round(mibs / plain_mibs, 1),
                  "batch_of_64_superchunks": dict(
                      {"value": round(n64 * super_bytes / b64_s / MIB, 2), "unit": "MiB/s",
                       "kernel_ms": round(ms64, 1),
                       "int32_valu_frac": round(n64 * info.nodes * OPS_PER_NODE / (ms64 / 1e3) / peak, 4),
                       "note": f"latency-bound: 64 graphs of {info.levels} levels"},
                      **b64_shape),
                  "streamed_64_superchunks": streamed,
                  "uneven_sessions_one_graph": {
                      "value": round(u_mibs, 2), "unit": "MiB/s", "replay_ms": round(u_ms, 1),
                      "steps": args.uneven_steps,
                      "int32_valu_frac": round(u_nodes * OPS_PER_NODE / (u_ms / 1e3) / peak, 4),
                      "workload": "48 sessions (PRNG seed 3), chunk_blocks in {2^12,2^14,2^15,2^17}, 1..32 "
                                  "super-chunks each, autotuned launch groups in one hipGraph; last "
                                  "super-chunk of every session byte-checked",
                      "super_chunks_by_cb": u_detail["super_chunks_by_cb"],
                      "autotune_s": u_detail["autotune_s"], "autotuned": u_detail["autotuned"]},
                  "serving_poisson": {
                      "unit": "MiB/s and ms per super-chunk (submit -> sink)",
                      "seconds_per_rate": args.extra_serve_seconds, "per_rate": s_results,
                      "checked_requests": s_checked,
                      "workload": "Poisson arrivals (PRNG seed 4), chunk_blocks 2^12..2^17 log-uniform, "
                                  "1..128 super-chunks per request, dry-run sink"}}
    topo.close()

    bpn = TRAFFIC_BYTES_PER_NODE.get((cb, info.lanes, args.mode, args.waves_per_simd))
    traffic = round(bpn * nodes) if bpn else None
    cpu = None
    if not args.no_cpu:
        cdt, kind = cpu_baseline(cb, key, rbo, args.cpu_sample, args.cpu_threads)
        log(f"cpu baseline ({kind}): {args.cpu_sample} super-chunks in {cdt:.2f} s")
        cpu = {"value": round(args.cpu_sample * super_bytes / cdt / MIB, 3), "unit": "MiB/s",
               "cores": args.cpu_threads, "kind": kind,
               "sample": f"{args.cpu_sample} super-chunks at chunk_blocks={cb} from the same start "
                         f"block, {cdt:.2f} s wall ({os.cpu_count()} host cpus visible, {cpu_model()})"}
        # The reference's default worker count (POSE_REGION_THREADS = 4) on a
        # quarter of the sample, for comparison with the all-cores figure.
        n4 = max(1, args.cpu_sample // 4)
        cdt4, kind4 = cpu_baseline(cb, key, rbo, n4, 4)
        log(f"cpu baseline ({kind4}, 4 threads): {n4} super-chunks in {cdt4:.2f} s")
        cpu["threads_4"] = {"value": round(n4 * super_bytes / cdt4 / MIB, 3), "unit": "MiB/s",
                            "kind": kind4, "sample": f"{n4} super-chunks, {cdt4:.2f} s wall"}
        tb = reference_topology_build(cb)
        if tb:
            log(f"reference topology (host): edges {tb['edges_s']:.2f} s + slot map {tb['slot_map_s']:.2f} s; "
                f"this build {build_s:.2f} s including upload")
            cpu["topology_build"] = dict(tb, this_build_s=round(build_s, 2))

    line = {
        "metric": "label MiB/s per MI355X at chunk_blocks=2^17 (batched persistent)",
        "value": round(mibs, 2),
        "unit": "MiB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (SplitMix64 key of PRNG seed 2, aligned random start)",
        "config": {"workload": f"batched persistent labeling, chunk_blocks={cb}, {n_super} "
                               f"super-chunks per launch (hipGraph replay) from block {rbo}",
                   "chunk_blocks": cb, "super_chunks": n_super, "lanes": info.lanes,
                   "schedule": "list" if args.mode == 1 else "levels",
                   "schedule_steps": info.steps, "levels": info.levels, "slots": pi.slots,
                   "lanes_per_node": pi.lanes_per_node,
                   "topology_build_s": round(build_s, 2), "parity_sample": js,
                   "autotuned": tuned},
        "roofline": {"bound": "valu", "achieved": round(achieved / 1e12, 3),
                     "peak": round(peak / 1e12, 2), "unit": "Tops/s (int32)",
                     "frac": round(achieved / peak, 4), "traffic": traffic,
                     "ops_per_node": OPS_PER_NODE, "kernel_ms": round(kernel_s * 1e3, 2)},
        "cpu_baseline": cpu,
    }
    line.update(extras)
    return line


def streamed_batch_through_server(args, pg, cus):
    """64 super-chunks at 2^17 (PRNG seed 2 key, random aligned start) through the
    label server: labels stream device -> pinned host -> dry-run sink; 16 randomly
    chosen super-chunks are recomputed on the CPU and byte-compared."""
    sm = SplitMix(2)
    key = sm.key()
    cb, n = 1 << 17, 64
    rbo = sm.aligned_start(cb)
    host = bytearray(n * cb * 32)
    # The host copy is the caller's buffer: fault its pages in before the
    # timed region, as a resident buffer would be.
    import numpy as np
    np.frombuffer(host, dtype=np.uint8)[::4096] = 1
    srv = pg.Server(hbm_budget_bytes=16 << 30, pinned_bytes=2 << 30, preload=[cb])
    t0 = time.perf_counter()
    srv.submit(key, cb, rbo, n, host_out=host)
    srv.drain()
    wall = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    picks = set()
    while len(picks) < 16:
        picks.add(sm.below(n))
    picks = sorted(picks)
    from pyoracle import Oracle
    exp = Oracle().label_superchunks(key, cb, picks, rbo, args.cpu_threads)
    sb = cb * 32
    for k, j in enumerate(picks):
        if bytes(host[j * sb:(j + 1) * sb]) != exp[k * sb:(k + 1) * sb]:
            print(f"PARITY FAILURE in streamed super-chunk {j}", file=sys.stderr)
            sys.exit(1)
    log(f"config 3 streamed: {n} super-chunks in {wall:.2f} s, {len(picks)} random super-chunks verified")
    return {"value": round(n * sb / wall / MIB, 2), "unit": "MiB/s", "super_chunks": n,
            "start_block": rbo, "p50_ms": round(st["p50_ms"], 1), "p99_ms": round(st["p99_ms"], 1),
            "kernel_ms": round(st["kernel_ms_total"], 1), "copy_ms": round(st["copy_ms_total"], 1),
            "sink_ms": round(st["sink_ms_total"], 1), "verified_random_superchunks": picks,
            "note": "device -> pinned host -> dry-run sink through the label server (latency form)"}


def workload_plain(args, torch, pg, dev, peak, cus):
    """Config: chunk_blocks=2^14, PRNG seed 1 key, 256 super-chunks from position 0, one
    super-chunk graph per launch, full byte compare against the CPU labeler."""
    sm = SplitMix(1)
    key = sm.key()
    cb, n = args.cb if args.cb != (1 << 17) else 1 << 14, args.super or 256
    sb = cb * 32
    topo = pg.Topology(cb, args.lanes, args.mode)
    out = torch.zeros(n * sb, dtype=torch.uint8, device=dev)
    # One untimed launch: the first one that labels the scaffold copies on
    # their own uploads their schedules (once per topology).
    pg.label_batch([(topo, key, 0, 1, out.data_ptr())], 1)
    torch.cuda.synchronize()
    kms = []
    t0 = time.perf_counter()
    for j in range(n):
        kms.append(pg.label_batch([(topo, key, j * cb, 1, out.data_ptr() + j * sb)], 1))
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    got = out.cpu().numpy().tobytes()
    exp = cpu_expected(key, cb, range(n), 0, args.cpu_threads)
    if got != exp:
        print("PARITY FAILURE: plain labels differ from the CPU labeler", file=sys.stderr)
        sys.exit(1)
    log(f"plain: {n} launches, full byte compare OK")
    kernel_s = sum(kms) / 1e3
    return {"metric": f"label MiB/s, plain one-graph-per-launch form, chunk_blocks={cb}",
            "value": round(n * sb / el / MIB, 3), "unit": "MiB/s", "n_gpus": 1, "steps": n,
            "warmup": 0, "ms_per_step": round(el / n * 1e3, 2), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "u32",
            "data": "synthetic (SplitMix64 key of PRNG seed 1)",
            "config": {"workload": f"{n} super-chunks from position 0, one launch each",
                       "chunk_blocks": cb, "super_chunks": n, "full_compare": True},
            "roofline": {"bound": "valu", "achieved": round(topo.info.nodes * n * OPS_PER_NODE / kernel_s / 1e12, 3),
                         "peak": round(peak / 1e12, 2), "unit": "Tops/s (int32)",
                         "frac": round(topo.info.nodes * n * OPS_PER_NODE / kernel_s / peak, 5),
                         "traffic": None},
            "cpu_baseline": None}


def measure_uneven(args, torch, pg, dev, steps, warmup):
    """48 sessions, independent keys from the PRNG-seed-3 stream; chunk_blocks uniform
    over {2^12, 2^14, 2^15, 2^17}; 1..32 super-chunks each at aligned random starts.
    The autotuner picks the launch shape (nodes per step, schedule, lanes per node,
    split, occupancy) per chunk_blocks for that chunk_blocks' share of the batch;
    sessions of equal shape share one persistent launch and all launches are captured
    in one hipGraph (parallel branches).  Sampled byte check: one super-chunk of
    every session.  Returns (MiB/s, mean replay ms, elapsed s, nodes per replay,
    detail dict)."""
    meta = uneven_sessions()
    topos, tuned = {}, {}
    total = sum(n * cb * 32 for _, cb, n, _ in meta)
    t_tune = time.perf_counter()
    for cb in sorted({m[1] for m in meta}):
        if args.no_tune:
            topos[cb] = pg.Topology(cb, args.lanes, args.mode)
            continue
        n_cb = sum(m[2] for m in meta if m[1] == cb)
        t0 = time.perf_counter()
        best, _ = pg.autotune(cb, n_cb)
        tuned[cb] = best
        log(f"autotune cb={cb} ({n_cb} super-chunks, {time.perf_counter() - t0:.1f} s): lanes {best['lanes']} "
            f"{'list' if best['mode'] == 1 else 'levels'} lanes/node {best['lanes_per_node']} "
            f"split {best['split']} wps {best['waves_per_simd']} slots {best['slots']}")
        topos[cb] = pg.Topology(cb, best["lanes"], best["mode"])
    t_tune = time.perf_counter() - t_tune
    outs = [torch.empty(n * cb * 32, dtype=torch.uint8, device=dev) for _, cb, n, _ in meta]
    sessions = [(topos[cb], key, rbo, n, o.data_ptr()) for (key, cb, n, rbo), o in zip(meta, outs)]
    if tuned:
        launches = [(tuned[cb]["slots"], tuned[cb]["waves_per_simd"], tuned[cb]["lanes_per_node"],
                     tuned[cb]["split"]) for _, cb, _, _ in meta]
        plan = pg.Plan(sessions, graph=True, launches=launches)
    else:
        plan = pg.Plan(sessions, args.slots, args.waves_per_simd, graph=True)
    groups = plan.groups()
    log(f"plan: {plan.info.groups} launch group(s) in one hipGraph: " +
        "; ".join(f"{g['super_chunks']} super-chunks, lanes {g['lanes']}, lanes/node {g['lanes_per_node']}, "
                  f"split {g['split']}, slots {g['slots']}" for g in groups))
    for _ in range(warmup):
        plan.run()
    torch.cuda.synchronize()
    kms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        kms.append(plan.run())
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    nodes = plan.info.nodes
    plan.close()
    # Sampled byte check: the last super-chunk of every session.
    exp = cpu_expected_many([(key, cb, rbo + (n - 1) * cb) for key, cb, n, rbo in meta], args.cpu_threads)
    for (key, cb, n, rbo), o, e in zip(meta, outs, exp):
        if o[(n - 1) * cb * 32:].cpu().numpy().tobytes() != e:
            print(f"PARITY FAILURE in session cb={cb} n={n}", file=sys.stderr)
            sys.exit(1)
    log("uneven: all 48 sessions' sampled super-chunks byte-identical")
    for t in topos.values():
        t.close()
    del outs
    mix = {str(cb): sum(m[2] for m in meta if m[1] == cb) for cb in topos}
    detail = {"super_chunks_by_cb": mix, "bytes_per_step": total,
              "autotuned": {str(cb): {k: v[k] for k in ("lanes", "mode", "lanes_per_node", "split",
                                                         "waves_per_simd", "slots")}
                            for cb, v in tuned.items()} or None,
              "autotune_s": round(t_tune, 1), "launch_groups": groups}
    return total * steps / el / MIB, sum(kms) / len(kms), el, nodes, detail


def workload_uneven(args, torch, pg, dev, peak, cus):
    mibs, kms, el, nodes, detail = measure_uneven(args, torch, pg, dev, args.steps, args.warmup)
    kernel_s = kms / 1e3
    return {"metric": "label MiB/s, 48 uneven sessions in one hipGraph-captured launch",
            "value": round(mibs, 2), "unit": "MiB/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(el / args.steps * 1e3, 2),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
            "data": "synthetic (SplitMix64 keys of PRNG seed 3)",
            "config": dict({"workload": "48 sessions, chunk_blocks in {2^12,2^14,2^15,2^17}, 1..32 "
                                        "super-chunks each"}, **detail),
            "roofline": {"bound": "valu", "achieved": round(nodes * OPS_PER_NODE / kernel_s / 1e12, 3),
                         "peak": round(peak / 1e12, 2), "unit": "Tops/s (int32)",
                         "frac": round(nodes * OPS_PER_NODE / kernel_s / peak, 4), "traffic": None},
            "cpu_baseline": None}


def challenged_nodes(pg, chal):
    """Scaffold nodes the verifier labels for these challenges: one labeling per
    distinct (chunk seed, chunk_blocks); only the copy of the scaffold that holds
    the challenged ranks when they all fall in one copy (the copies share no
    edge), else the whole scaffold."""
    import ctypes as C
    by_chunk, topo_parts = {}, {}
    for seed, cb, rank in chal:
        by_chunk.setdefault((bytes(seed), cb), set()).add(rank)
    nodes = 0
    for (seed, cb), ranks in by_chunk.items():
        if cb not in topo_parts:
            ht = pg.Topology(cb, host_only=True)
            parts = []
            for c in (0, 1):
                info = pg.TopoPart()
                if pg.lib.pose_gpu_topo_part(ht.h, c, C.byref(info), None, None, None) == 0:
                    parts.append(info)
            topo_parts[cb] = (ht.info.nodes, parts)
            ht.close()
        whole, parts = topo_parts[cb]
        copies = {0 if parts and r < parts[0].end_rank else 1 for r in ranks} if parts else {0, 1}
        if len(copies) == 1:
            p = parts[copies.pop()]
            nodes += p.end_node - p.first_node
        else:
            nodes += whole
    return nodes


def workload_verify(args, torch, pg, dev, peak, cus):
    """Challenge answers at chunk_blocks=2^17 (PRNG seed 5): one session key,
    `--challenges` challenges on distinct random super-chunks of a 2^40-byte
    region, each at a random rank.  Every answer labels its whole super-chunk
    graph (36,175,870 node hashes).  The reference's CPU pose_graph_challenge
    answers a bounded sample on one core, and its answers are compared."""
    sm = SplitMix(5)
    session = sm.key()
    cb = args.cb
    n_chunks = (1 << 40) // (cb * 32)
    idxs = set()
    while len(idxs) < args.challenges:
        idxs.add(sm.below(n_chunks))
    chal = [(pg.chunk_seed(session, i), cb, sm.below(cb)) for i in sorted(idxs)]
    rates = {}
    for n in sorted({min(64, len(chal)), len(chal)}):
        # Untimed warmup of the same batch shape: the first launch of a shape on
        # a new topology runs up to 1.7x slower (1,024 super-chunks of 2^17:
        # 2.99 s first, 1.75 s after).
        pg.challenge_batch(chal[:n])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        got, ms = pg.challenge_batch(chal[:n], with_time=True)
        el = time.perf_counter() - t0
        rates[n] = (n / el, el, ms)
        log(f"verify: {n} challenges in {el:.2f} s (kernel {ms:.0f} ms)")
    from pyoracle import RefLib, ref_available
    cpu = None
    if not args.no_cpu and ref_available(4):
        ref = RefLib(4)
        k = max(1, args.cpu_verify_sample)
        t0 = time.perf_counter()
        exp = [ref.challenge(s, c, r) for s, c, r in chal[:k]]
        cdt = time.perf_counter() - t0
        ref.close()
        if exp != got[:k]:
            print("PARITY FAILURE: challenge answers differ from the reference", file=sys.stderr)
            sys.exit(1)
        log(f"cpu pose_graph_challenge: {k} answers in {cdt:.2f} s, identical")
        cpu = {"value": round(k / cdt, 4), "unit": "answers/s", "cores": 1, "kind": "reference",
               "sample": f"{k} challenges via the reference's pose_graph_challenge, {cdt:.2f} s "
                         f"({cpu_model()})"}
    n_all = len(chal)
    value, el, ms = rates[n_all]
    nodes = challenged_nodes(pg, chal)
    return {"metric": f"challenge answers/s at chunk_blocks=2^{cb.bit_length() - 1} (full super-chunk "
                      "labeling per answer)",
            "value": round(value, 2), "unit": "answers/s", "n_gpus": 1, "steps": 1, "warmup": 1,
            "ms_per_step": round(el * 1e3, 1), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "u32", "data": "synthetic (SplitMix64 key of PRNG seed 5)",
            "config": {"workload": f"{n_all} challenges on distinct super-chunks, random ranks",
                       "chunk_blocks": cb,
                       "by_batch": {str(n): {"answers_per_s": round(r[0], 2), "wall_s": round(r[1], 3),
                                             "kernel_ms": round(r[2], 1)} for n, r in rates.items()}},
            "roofline": {"bound": "valu", "achieved": round(nodes * OPS_PER_NODE / (ms / 1e3) / 1e12, 3),
                         "peak": round(peak / 1e12, 2), "unit": "Tops/s (int32)",
                         "frac": round(nodes * OPS_PER_NODE / (ms / 1e3) / peak, 4), "traffic": None},
            "cpu_baseline": cpu}


def measure_serving(args, pg, rates, seconds):
    """Poisson request streams (PRNG seed 4, one stream across the rates): chunk_blocks
    log-uniform over 2^12..2^17, 1..128 super-chunks, aligned random starts, own key
    per request; dry-run sink; a fresh server per rate, so its statistics cover that
    rate only.  Up to six small requests are kept and byte-checked on the CPU."""
    results, samples, sm = [], [], None
    for lam in rates:
        srv = pg.Server(hbm_budget_bytes=int(args.hbm_budget_gib * 2**30),
                        pinned_bytes=int(args.pinned_gib * 2**30),
                        max_super_per_launch=args.max_super_per_launch,
                        preload=[1 << k for k in range(12, 18)])
        reqs, sm = poisson_requests(lam, seconds, sm=sm)
        t0 = time.perf_counter()
        for t, key, cb, n, rbo in reqs:
            now = time.perf_counter() - t0
            if now < t:
                time.sleep(t - now)
            keep = None
            if len(samples) < 6 and n <= 4:
                keep = bytearray(n * cb * 32)
                samples.append((key, cb, n, rbo, keep))
            srv.submit(key, cb, rbo, n, host_out=keep)
        if srv.drain() != 0:
            print(f"SERVING FAILURE at {lam} requests/s", file=sys.stderr)
            sys.exit(1)
        st = srv.stats()
        srv.close()
        log(f"rate {lam}/s: {len(reqs)} requests, {st['super_chunks_done']} super-chunks, "
            f"p50 {st['p50_ms']:.0f} ms p99 {st['p99_ms']:.0f} ms, {st['mib_per_s']:.0f} MiB/s")
        r = {"rate_per_s": lam, "requests": len(reqs), "super_chunks": st["super_chunks_done"]}
        r.update({k: round(st[k], 2) for k in ("p50_ms", "p99_ms", "max_ms", "mib_per_s")})
        r.update({"peak_inflight": st["peak_inflight"], "launches": st["launches"]})
        results.append(r)
    items = [(key, cb, rbo + j * cb) for key, cb, n, rbo, _ in samples for j in range(n)]
    exp = iter(cpu_expected_many(items, args.cpu_threads))
    for key, cb, n, rbo, buf in samples:
        for j in range(n):
            if bytes(buf[j * cb * 32:(j + 1) * cb * 32]) != next(exp):
                print("PARITY FAILURE in served request", file=sys.stderr)
                sys.exit(1)
    log(f"serve: {len(samples)} sampled requests byte-identical to the CPU labeler")
    return results, len(samples)


def workload_serve(args, torch, pg, dev, peak, cus):
    rates = [float(r) for r in args.rates.split(",")]
    results, n_checked = measure_serving(args, pg, rates, args.serve_seconds)
    best = max(r["mib_per_s"] for r in results)
    return {"metric": "served label MiB/s and p50/p99 per-super-chunk latency (Poisson arrivals)",
            "value": best, "unit": "MiB/s", "n_gpus": 1, "steps": len(rates), "warmup": 0,
            "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "u32", "data": "synthetic (SplitMix64 keys of PRNG seed 4)",
            "config": {"workload": "continuous-batching server, dry-run sink, chunk_blocks "
                                   "2^12..2^17 log-uniform, 1..128 super-chunks per request",
                       "seconds_per_rate": args.serve_seconds, "per_rate": results,
                       "checked_requests": n_checked},
            "roofline": None, "cpu_baseline": None}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="batched", choices=["batched", "plain", "uneven", "serve", "verify"])
    ap.add_argument("--cb", type=int, default=1 << 17)
    ap.add_argument("--super", type=int, default=0, help="super-chunks per batch (0 = workload default)")
    ap.add_argument("--lanes", type=int, default=64)
    ap.add_argument("--mode", type=int, default=1)
    ap.add_argument("--slots", type=int, default=0)
    ap.add_argument("--waves-per-simd", type=int, default=4)
    ap.add_argument("--plain-launches", type=int, default=2)
    ap.add_argument("--cpu-sample", type=int, default=16)
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--rates", default="2,4,8,16,32,64")
    ap.add_argument("--serve-seconds", type=float, default=6.0)
    ap.add_argument("--hbm-budget-gib", type=float, default=160.0)
    ap.add_argument("--max-super-per-launch", type=int, default=0,
                    help="serve workload: label server's cap per launch (0 = its default)")
    ap.add_argument("--challenges", type=int, default=1024, help="verify workload: challenges per call")
    ap.add_argument("--cpu-verify-sample", type=int, default=2,
                    help="verify workload: challenges answered by the reference CPU path")
    ap.add_argument("--pinned-gib", type=float, default=32.0,
                    help="serve workload: pinned host staging ring")
    ap.add_argument("--lanes-per-node", type=lambda v: 0xFFFFFFFF if v == "auto" else int(v), default="auto",
                    help="1, 4 or auto (four-lane latency form for small batches)")
    ap.add_argument("--autotune", action="store_true", help="pick lanes/occupancy with pose_gpu_autotune")
    ap.add_argument("--no-tune", action="store_true",
                    help="uneven workload: one automatic shape for all sessions instead of the autotuner")
    ap.add_argument("--uneven-steps", type=int, default=2,
                    help="batched workload: timed replays of the uneven-sessions extra")
    ap.add_argument("--extra-rates", default="4,16,64",
                    help="batched workload: arrival rates of the serving extra (requests/s)")
    ap.add_argument("--extra-serve-seconds", type=float, default=4.0,
                    help="batched workload: seconds of arrivals per rate in the serving extra")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="run a copy kernel of this size first (counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU: super-chunks are independent graphs, so a multi-GPU "
              "run is N disjoint copies of this one (no collectives); run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import pose_gpu as pg

    dev = torch.device("cuda:0")
    props = torch.cuda.get_device_properties(dev)
    peak, cus = valu_peak(props)
    if args.copy_calibration_mib:
        # A plain device copy of known size (int32 x + 1: reads N bytes, writes
        # N bytes) so a counter pass can scale FETCH_SIZE / WRITE_SIZE.
        n = args.copy_calibration_mib * MIB // 4
        x = torch.zeros(n, dtype=torch.int32, device=dev)
        y = x + 1
        torch.cuda.synchronize()
        log(f"copy calibration: {args.copy_calibration_mib} MiB read + written")
        del x, y
    fn = {"batched": workload_batched, "plain": workload_plain, "uneven": workload_uneven,
          "serve": workload_serve, "verify": workload_verify}[args.workload]
    print(json.dumps(fn(args, torch, pg, dev, peak, cus)))


if __name__ == "__main__":
    main()
