import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "pose-memory-wiping_amd")
for p in (REPO, PKG, os.path.join(REPO, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: long CPU-side check")


@pytest.fixture(scope="session")
def golden():
    import json
    with open(os.path.join(REPO, "tests", "golden", "golden.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def oracle():
    from pyoracle import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def cpu_superchunks():
    """expected(items) -> list of labels: items are (session seed, chunk_blocks,
    absolute block of the super-chunk); each super-chunk is labeled on the CPU
    by the reference build (when present; the oracle otherwise), many at once
    on a thread pool (the C libraries run without the GIL)."""
    import threading
    from concurrent.futures import ThreadPoolExecutor
    from pyoracle import Oracle, RefLib, ref_available

    local = threading.local()
    made = []
    lock = threading.Lock()
    use_ref = ref_available(4)

    def one(item):
        seed, cb, block = item
        if use_ref:
            if not hasattr(local, "ref"):
                local.ref = RefLib(4)
                with lock:
                    made.append(local.ref)
            return local.ref.label_region(seed, cb, 1, block)
        if not hasattr(local, "oracle"):
            local.oracle = Oracle()
        return local.oracle.label_superchunks(seed, cb, [0], block, 1)

    ex = ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1))

    def expected(items):
        return list(ex.map(one, list(items)))

    yield expected
    ex.shutdown(wait=True)
    for r in made:
        r.close()
