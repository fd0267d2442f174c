"""The C++ facade (xpir_amd/cpp/libpir_amd.hpp, the reference's libpir.hpp server API)
driven end to end by tests/cpp/simple_pir.cpp, compared with the oracle."""
import os
import subprocess

import numpy as np
import pytest

from oracle.oracle import Oracle, extract_reply

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEMO = os.path.join(ROOT, "xpir_amd", "_build", "simple_pir")


@pytest.mark.parametrize("crypto,n,L,alpha,dims", [
    ("LWE:80:1024:60", 50, 3000, 1, [50]),
    ("LWE:80:2048:120", 30, 4096, 1, [6, 5]),
    ("LWE:80:2048:120", 12, 1000, 3, [4]),
    ("LWE:80:2048:120", 13, 1000, 3, [5]),   # partial last aggregate: the reference's bytes
])
def test_simple_pir_through_the_facade(tmp_path, crypto, n, L, alpha, dims):
    out = tmp_path / "reply.bin"
    index = -(-n // alpha) - 1
    args = [DEMO, crypto, str(n), str(L), str(alpha), str(len(dims)), *map(str, dims), "1", str(index), "3", "4",
            str(out)]
    print("run:", " ".join(args[1:]), flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    got = np.fromfile(out, dtype=np.uint64)
    orc = Oracle(crypto, dims[0])
    data = orc.rand_bytes(1, 0, n * L)
    unit = L * alpha
    units = -(-n // alpha)
    padded = Oracle.aggregate_units(data, n, L, alpha).reshape(-1)
    P = orc.polys(unit)
    words = np.zeros(int(np.prod(dims)) * P * orc.mn, dtype=np.uint64)
    words[:units * P * orc.mn] = np.concatenate([orc.import_element(padded[u * unit:(u + 1) * unit])
                                                  for u in range(units)])
    sk = orc.keygen(4)
    q = orc.make_query(sk, dims, index, 3)
    want, _ = orc.generate_reply(words, P, dims, q)
    assert np.array_equal(got, want)
    assert extract_reply(orc, sk, got, dims, unit) == padded[index * unit:(index + 1) * unit].tobytes()


def test_simple_pir_from_a_directory(tmp_path):
    """The same through DBDirectoryProcessor: files of uneven length in directory-listing
    order, shorter ones zero-padded to the largest."""
    crypto, n, L, dims = "LWE:80:2048:120", 30, 4096, [6, 5]
    out = tmp_path / "reply.bin"
    index = 17
    args = [DEMO, crypto, str(n), str(L), "1", str(len(dims)), *map(str, dims), "1", str(index), "3", "4",
            str(out), str(tmp_path / "db")]
    print("run:", " ".join(args[1:]), flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    lines = (tmp_path / "reply.bin.catalog").read_text().split("\n")
    assert lines[0] == "0" and int(lines[1]) == n
    names, sizes = lines[2:2 + 2 * n:2], [int(s) for s in lines[3:3 + 2 * n:2]]
    orc = Oracle(crypto, dims[0])
    data = orc.rand_bytes(1, 0, n * L)
    elems = []
    for name, size in zip(names, sizes):
        e = int(name[1:])
        assert size == L - e % 3
        elems.append(np.concatenate([data[e * L:e * L + size], np.zeros(L - size, dtype=np.uint8)]))
    P = orc.polys(L)
    words = np.concatenate([orc.import_element(x) for x in elems])
    sk = orc.keygen(4)
    q = orc.make_query(sk, dims, index, 3)
    want, _ = orc.generate_reply(words, P, dims, q)
    got = np.fromfile(out, dtype=np.uint64)
    assert np.array_equal(got, want)
    assert extract_reply(orc, sk, got, dims, L) == elems[index].tobytes()


def test_pushed_query_staging(tmp_path):
    """PIRReplyGenerator::pushQuery stages elements through pinned chunks into HBM: a query
    over three chunks, a second reply, a new query after freeQueries, two levels, and
    element counts off by one (tests/cpp/facade_staging.cpp), against the C ABI's reply."""
    exe = tmp_path / "facade_staging"
    pkg = os.path.join(ROOT, "xpir_amd")
    subprocess.run(["g++", "-O1", "-std=c++17", "-Wall", "-o", str(exe),
                    os.path.join(ROOT, "tests", "cpp", "facade_staging.cpp"),
                    "-L", pkg, "-lpir_amd", "-lxpir_amd", f"-Wl,-rpath,{pkg}"], check=True)
    print("run: facade_staging", flush=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stderr + r.stdout


def test_client_classes_round_trip(tmp_path):
    """PIRQueryGenerator and PIRReplyExtraction of the facade (on the GPU) around
    PIRReplyGenerator, as simplePIR uses them: d = 1 and 2, alpha = 3 with a partial last
    aggregate, uneven elements; reproducible seeds; refused inputs
    (tests/cpp/client_classes.cpp)."""
    exe = tmp_path / "client_classes"
    pkg = os.path.join(ROOT, "xpir_amd")
    subprocess.run(["g++", "-O1", "-std=c++17", "-Wall", "-o", str(exe),
                    os.path.join(ROOT, "tests", "cpp", "client_classes.cpp"),
                    "-L", pkg, "-lpir_amd", "-lxpir_amd", f"-Wl,-rpath,{pkg}"], check=True)
    print("run: client_classes", flush=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stderr + r.stdout
