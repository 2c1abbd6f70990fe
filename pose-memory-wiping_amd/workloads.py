"""Synthetic workloads of the labeling benchmarks and tests.

Every draw comes from one SplitMix64 stream per workload (Steele, Lea and
Flood 2014): a 32-byte session key is four outputs, little-endian; start
positions, chunk_blocks, super-chunk counts and Poisson gaps are later draws
of the same stream.  A position is "aligned" when it is a whole number of
super-chunks (chunk_blocks x 32 bytes); starts are drawn in [0, 2^40) bytes.
"""
import math


class SplitMix:
    """SplitMix64 stream; seed 0 gives 0xe220a8397b1dcdaf first."""

    def __init__(self, seed):
        self.s = seed & 0xFFFFFFFFFFFFFFFF

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        return z ^ (z >> 31)

    def key(self):
        return b"".join(self.next().to_bytes(8, "little") for _ in range(4))

    def uniform(self):
        """Uniform in [0, 1) from the top 53 bits."""
        return (self.next() >> 11) * (1.0 / (1 << 53))

    def below(self, n):
        return self.next() % n

    def aligned_start(self, cb):
        """Aligned absolute block offset of a position in [0, 2^40) bytes."""
        return self.below((1 << 40) // (cb * 32)) * cb


def streamed_batch(prng_seed=2, chunk_blocks=1 << 17, n_super=64):
    """One session of n_super consecutive super-chunks at a random aligned
    start: (key, chunk_blocks, n_super, start block, stream for later draws)."""
    sm = SplitMix(prng_seed)
    key = sm.key()
    return key, chunk_blocks, n_super, sm.aligned_start(chunk_blocks), sm


def uneven_sessions(prng_seed=3, n_sessions=48, sizes=(1 << 12, 1 << 14, 1 << 15, 1 << 17), max_super=32):
    """Sessions of independent keys: chunk_blocks uniform over `sizes`,
    1..max_super super-chunks, aligned random starts.  Returns a list of
    (key, chunk_blocks, n_super, start block)."""
    sm = SplitMix(prng_seed)
    out = []
    for _ in range(n_sessions):
        key = sm.key()
        cb = sizes[sm.below(len(sizes))]
        n = 1 + sm.below(max_super)
        out.append((key, cb, n, sm.aligned_start(cb)))
    return out


def poisson_requests(rate_per_s, seconds, prng_seed=4, sm=None, log2_cb=(12, 17), max_super=128):
    """Label requests of a Poisson stream: (arrival time in s from 0, key,
    chunk_blocks, n_super, start block).  chunk_blocks is log-uniform over
    2^log2_cb[0] .. 2^log2_cb[1] (powers of two), n_super uniform in
    1..max_super.  Pass `sm` to continue an existing stream."""
    sm = sm or SplitMix(prng_seed)
    lo, hi = log2_cb
    t, out = 0.0, []
    while t < seconds:
        key = sm.key()
        cb = 1 << (lo + sm.below(hi - lo + 1))
        n = 1 + sm.below(max_super)
        out.append((t, key, cb, n, sm.aligned_start(cb)))
        t += -math.log(1.0 - sm.uniform()) / rate_per_s
    return out, sm
