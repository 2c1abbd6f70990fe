"""pose_graph_hbm_topo_label_noalloc (fixed device scratch, no allocation)
against pose_graph_hbm_topo_label on the same inputs: byte-identical labels
and untouched tails, on two calls in a row, the second with a smaller input.
Cases cover the four-lane form whole and split into copies (including more
copy jobs than the fixed scratch holds slot regions for, at 2^17), the
one-lane form over more job records than one window holds, a 128-lane
topology and a chunk_blocks that is not a power of two."""
import ctypes as C

import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import pose_gpu as pg  # noqa: E402

_f = pg.lib.pose_graph_hbm_topo_label_noalloc
_f.argtypes = [C.c_void_p, C.c_char_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_uint64]
_f.restype = C.c_int


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")
    return torch.device("cuda:0")


# (chunk_blocks, topology lanes, super-chunks of the first call, of the second)
CASES = [(1 << 10, 64, 300, 7), (256, 64, 2000, 900), (1 << 12, 64, 40, 3), (1 << 17, 64, 40, 1),
         (1000, 64, 5, 2), (1 << 11, 128, 50, 4)]


@pytest.mark.parametrize("cb,lanes,n1,n2", CASES)
def test_noalloc_equals_topo_label_twice(dev, oracle, cb, lanes, n1, n2):
    T = pg.Topology(cb, lanes, 1)
    sb = cb * 32
    for call, (n, rbo, seed) in enumerate(((n1, 5 * cb, b"\x5a" * 32), (n2, 77 * cb, b"\x3c" * 19))):
        length = n * sb + 100          # a tail shorter than a super-chunk stays untouched
        say(f"cb={cb} lanes={lanes} call {call}: {n} super-chunks")
        a = torch.full((length,), 0xA5, dtype=torch.uint8, device=dev)
        b = torch.full((length,), 0xA5, dtype=torch.uint8, device=dev)
        assert T.label(seed, a.data_ptr(), length, rbo) == 0
        assert _f(T.h, seed, len(seed), C.c_void_p(b.data_ptr()), length, rbo) == 0
        torch.cuda.synchronize()
        ha, hb = a.cpu().numpy().tobytes(), b.cpu().numpy().tobytes()
        assert ha[n * sb:] == b"\xa5" * 100
        assert ha == hb, (cb, call)
        if cb <= 1 << 12:
            assert hb[:sb] == oracle.label_superchunks(seed, cb, [0], rbo, 4)
    T.close()


def test_noalloc_argument_errors(dev):
    T = pg.Topology(1 << 8)
    buf = torch.zeros(1 << 13, dtype=torch.uint8, device=dev)
    assert _f(None, b"k" * 32, 32, C.c_void_p(buf.data_ptr()), 1 << 13, 0) == -1
    assert _f(T.h, b"k" * 32, 32, None, 1 << 13, 0) == -1
    # shorter than one super-chunk: nothing to label, like the reference
    assert _f(T.h, b"k" * 32, 32, C.c_void_p(buf.data_ptr()), (1 << 8) * 32 - 1, 0) == 0
    T.close()
