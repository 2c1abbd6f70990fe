#!/usr/bin/env python3
"""Label-rate benchmark of the MI355X PoSE-DB labeler.

One "step" is one pass of the hot path over one batch: a single persistent
launch (pose_gpu_label_batch) that labels `--super` super-chunks of
chunk_blocks = 2^17 (keyed BLAKE3, exact integer arithmetic) into a device
buffer already resident in HBM.  The session key is the PRNG-seed-2 SplitMix64
key and the batch starts at an aligned absolute position drawn from the same
stream, as in the 2^17 batched workload.  W warmup steps, then K timed steps
between device synchronisations; the JSON line reports label MiB/s over the K
steps.

Every run re-checks sampled super-chunks of the GPU output against the CPU
labeler (byte compare) and times the reference CPU labeler (built from source
under oracle/_ref) on a bounded sample of the same workload on the host.
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "pose-memory-wiping_amd"))
sys.path.insert(0, os.path.join(REPO, "oracle"))

# int32 VALU ops executed per scaffold node by the labeling kernel: two BLAKE3
# compressions (block 1 and block 2), 7 rounds x 8 G x 12 instructions + 8
# output XORs each.  Block 0 is part of the precomputed topology.
OPS_PER_NODE = 2 * (7 * 8 * 12 + 8)
# The full node hash (three compressions), used for the equivalent rate.
OPS_PER_NODE_FULL = 3 * (7 * 8 * 12 + 8)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def splitmix(state):
    state = (state + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return state, z ^ (z >> 31)


def workload_key(prng_seed):
    st = prng_seed
    key = b""
    for _ in range(4):
        st, v = splitmix(st)
        key += v.to_bytes(8, "little")
    return st, key


def valu_peak(props):
    cus = props.multi_processor_count
    # MI355X: 64 int32 lanes per CU per clock, 2.4 GHz peak engine clock.
    return cus * 64 * 2.4e9, cus


def cpu_baseline(cb, key, rbo, n_sample, threads):
    from pyoracle import RefLib, ref_available, Oracle
    if ref_available(threads):
        ref = RefLib(threads)
        t0 = time.perf_counter()
        out = ref.label_region(key, cb, n_sample, rbo)
        dt = time.perf_counter() - t0
        ref.close()
        kind = "reference"
    else:
        o = Oracle()
        t0 = time.perf_counter()
        out = o.label_region(key, cb, n_sample, rbo, threads)
        dt = time.perf_counter() - t0
        kind = "port"
    return out, dt, kind


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cb", type=int, default=1 << 17)
    ap.add_argument("--super", type=int, default=0, help="super-chunks per batch (0 = 2 per resident slot)")
    ap.add_argument("--lanes", type=int, default=64)
    ap.add_argument("--mode", type=int, default=1)
    ap.add_argument("--slots", type=int, default=0)
    ap.add_argument("--waves-per-simd", type=int, default=2)
    ap.add_argument("--verify", type=int, default=2, help="super-chunks re-checked on the CPU")
    ap.add_argument("--cpu-sample", type=int, default=16, help="super-chunks in the CPU baseline sample")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
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

    st, key = workload_key(2)
    st, draw = splitmix(st)
    cb = args.cb
    rbo = (draw % (1 << 18)) * (1 << 22) // 32          # aligned start in [0, 2^40) bytes
    rbo -= rbo % cb

    t0 = time.perf_counter()
    topo = pg.Topology(cb, args.lanes, args.mode)
    build_s = time.perf_counter() - t0
    info = topo.info
    log(f"topology cb={cb} lanes={info.lanes} steps={info.steps} levels={info.levels} "
        f"pool={info.pool_slots} build={build_s:.2f}s")

    slots = args.slots
    if args.super:
        n_super = args.super
    else:
        n_super = 2 * (slots or cus * max(1, (4 * args.waves_per_simd) // (args.lanes // 64)))
    super_bytes = cb * 32
    out = torch.empty(n_super * super_bytes, dtype=torch.uint8, device=dev)
    sess = [(topo, key, rbo, n_super, out.data_ptr())]

    for w in range(args.warmup):
        ms = pg.label_batch(sess, slots, args.waves_per_simd)
        log(f"warmup {w}: {n_super} super-chunks, kernel {ms:.1f} ms")
    torch.cuda.synchronize()
    kms = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        kms.append(pg.label_batch(sess, slots, args.waves_per_simd))
        log(f"step {k}: kernel {kms[-1]:.1f} ms")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0

    total_bytes = n_super * super_bytes * args.steps
    mibs = total_bytes / elapsed / (1 << 20)
    kernel_s = sum(kms) / len(kms) / 1e3
    nodes = info.nodes * n_super
    achieved = nodes * OPS_PER_NODE / kernel_s

    # Sampled CPU recompute of the GPU output (first and last super-chunk, then
    # evenly spaced ones).
    js = sorted({0, n_super - 1} | {int(i * n_super / max(1, args.verify)) for i in range(args.verify)})
    from pyoracle import Oracle
    o = Oracle()
    exp = o.label_superchunks(key, cb, js, rbo, args.cpu_threads)
    host = out.view(n_super, super_bytes)
    ok = all(host[j].cpu().numpy().tobytes() == exp[k * super_bytes:(k + 1) * super_bytes]
             for k, j in enumerate(js))
    if not ok:
        print("PARITY FAILURE: GPU labels differ from the CPU labeler", file=sys.stderr)
        sys.exit(1)
    log(f"parity: super-chunks {js} byte-identical to the CPU labeler")

    cpu = None
    if not args.no_cpu:
        _, cdt, kind = cpu_baseline(cb, key, rbo, args.cpu_sample, args.cpu_threads)
        log(f"cpu baseline ({kind}): {args.cpu_sample} super-chunks in {cdt:.2f} s")
        cpu = {"value": args.cpu_sample * super_bytes / cdt / (1 << 20), "unit": "MiB/s",
               "cores": args.cpu_threads, "kind": kind,
               "sample": f"{args.cpu_sample} super-chunks at chunk_blocks={cb} from the same start, "
                         f"{cdt:.2f} s wall, {os.cpu_count()} host cpus visible"}

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
        "data": "synthetic (SplitMix64 key, PRNG seed 2)",
        "config": {"workload": f"batched persistent labeling, chunk_blocks={cb}, {n_super} super-chunks "
                               f"per launch from block {rbo}", "chunk_blocks": cb, "super_chunks": n_super,
                   "lanes": info.lanes, "schedule_steps": info.steps, "levels": info.levels,
                   "slots": slots or "auto", "waves_per_simd": args.waves_per_simd,
                   "topology_build_s": round(build_s, 2), "parity_sample": js},
        "roofline": {"bound": "valu", "achieved": round(achieved / 1e12, 3), "peak": round(peak / 1e12, 2),
                     "unit": "Tops/s (int32)", "frac": round(achieved / peak, 4), "traffic": None,
                     "ops_per_node": OPS_PER_NODE, "kernel_ms": round(kernel_s * 1e3, 2),
                     "equiv_full_hash_frac": round(nodes * OPS_PER_NODE_FULL / kernel_s / peak, 4)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
