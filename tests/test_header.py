"""The C ABI header as a drop-in for the reference's common/include/gpu_label.h.

The reference producer includes gpu_label.h before graph.h and hash.h
(tools/disk-wipe-bench/disk_wipe_bench.cu:55-57); other callers include
hash.h first.  Both orders must compile, in C11 and C++17, against the
reference's real graph.h / hash.h, and the header must also stand alone.
"""
import os
import shutil
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
OUR_INC = os.path.join(REPO, "include")
REF_INC = "/root/reference/common/include"

ORDERS = {
    "producer order (gpu_label.h, graph.h, hash.h)": ["gpu_label.h", "graph.h", "hash.h"],
    "hash.h first": ["hash.h", "graph.h", "gpu_label.h"],
    "graph.h first": ["graph.h", "gpu_label.h", "hash.h"],
}

BODY = """
int main(void) {
    pose_hash_algo_t a = POSE_HASH_BLAKE3;
    void *p1 = (void *)&pose_graph_hbm_topo_build;       /* gpu_label.h */
    void *p2 = (void *)&pose_graph_label_region_pooled_inplace;   /* graph.h */
    (void)a; (void)p1; (void)p2;
    return POSE_HASH_SHA256 == 1 && POSE_HASH_AES_PRF == 2 ? 0 : 1;
}
"""


def _compile(tmp_path, name, includes, lang, inc_dirs, body=BODY):
    ext = "c" if lang == "c" else "cpp"
    src = tmp_path / f"{name}.{ext}"
    src.write_text("".join(f'#include "{h}"\n' for h in includes) + body)
    cc = ["gcc", "-std=c11"] if lang == "c" else ["g++", "-std=c++17"]
    cmd = cc + ["-Wall", "-Werror", "-fsyntax-only"] + [f"-I{d}" for d in inc_dirs] + [str(src)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return r.returncode, r.stderr


@pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("g++") is None, reason="needs gcc/g++")
@pytest.mark.skipif(not os.path.isdir(REF_INC),
                    reason="the reference checkout (/root/reference) is absent: no graph.h/hash.h to compile against")
@pytest.mark.parametrize("lang", ["c", "c++"])
@pytest.mark.parametrize("order", list(ORDERS))
def test_header_compiles_with_reference_headers(tmp_path, lang, order):
    rc, err = _compile(tmp_path, "tu", ORDERS[order], lang, [OUR_INC, REF_INC])
    assert rc == 0, err


@pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("g++") is None, reason="needs gcc/g++")
@pytest.mark.parametrize("lang", ["c", "c++"])
def test_header_stands_alone(tmp_path, lang):
    body = """
int main(void) {
    pose_hash_algo_t a = POSE_HASH_BLAKE3;
    (void)a;
    return (void *)&pose_graph_hbm_topo_build != 0 && POSE_HASH_AES_PRF == 2 ? 0 : 1;
}
"""
    rc, err = _compile(tmp_path, "alone", ["gpu_label.h"], lang, [OUR_INC], body)
    assert rc == 0, err


def test_binding_loads_one_hip_runtime():
    """The ctypes binding loads torch (when installed) before the library, so
    a process that imports the binding first still holds one HIP runtime:
    torch's copy, which the library's libamdhip64.so.7 then resolves to."""
    import ast
    import sys
    pytest.importorskip("torch")
    code = ("import sys; sys.path.insert(0, %r); import pose_gpu; import torch; "
            "m = open('/proc/self/maps').read(); "
            "print(sorted({l.split()[-1] for l in m.splitlines() if 'libamdhip64' in l}))"
            % os.path.join(REPO, "pose-memory-wiping_amd"))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr
    libs = ast.literal_eval(r.stdout.strip())
    assert len(libs) == 1, libs
