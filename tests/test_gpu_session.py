"""Whole sessions with the server program (xpir_amd/_build/pir_server, the reference's
PIRSession and PIRServer) on the GPU, driven by the reference client's side of the
protocol (xpir_amd/pir_client.py): the query goes up element by element, the reply comes
back ciphertext by ciphertext, equals the oracle's reply bytes and decrypts to the
chosen element.  Covers the three server modes (client-driven, -z with its saved
parameters and kept database, -L) and the import overlapped with the upload or not."""
import os
import struct
import sys

import numpy as np
import pytest

from oracle.oracle import Oracle, extract_reply

from xpir_amd.pir_client import CLIENT_DRIVEN, SERVER_DRIVEN

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from test_session import Server  # noqa: E402

pytestmark = pytest.mark.gpu


def _make_db(tmp_path, orc, n, L):
    """n files of L - (i % 3) bytes; returns the directory and name -> bytes"""
    db = tmp_path / "db"
    db.mkdir()
    data = orc.rand_bytes(1, 0, n * L)
    files = {}
    for i in range(n):
        b = data[i * L:i * L + L - i % 3].tobytes()
        name = f"f{i:03d}"
        (db / name).write_bytes(b)
        files[name] = b
    return db, files


def _expected(orc, files, names, maxb, alpha, dims, sk, index, seed):
    """the oracle's reply for the catalog's element order, and the query it answers"""
    unit = maxb * alpha
    units = -(-len(names) // alpha)
    padded = np.zeros(units * unit, dtype=np.uint8)
    for i, name in enumerate(names):
        b = np.frombuffer(files[name], dtype=np.uint8)
        padded[i * maxb:i * maxb + b.size] = b
    P = orc.polys(unit)
    words = np.zeros(int(np.prod(dims)) * P * orc.mn, dtype=np.uint64)
    words[:units * P * orc.mn] = np.concatenate([orc.import_element(padded[u * unit:(u + 1) * unit])
                                                 for u in range(units)])
    q = orc.make_query(sk, dims, index, seed)
    want, _ = orc.generate_reply(words, P, dims, q)
    return q, want, padded[index * unit:(index + 1) * unit].tobytes(), unit


def _exchange_query(c, orc, q, want):
    ct = 2 * orc.mn * 8
    qb = np.concatenate(q).tobytes()  # levels back to back, as the client uploads them
    c.send_query([qb[i:i + ct] for i in range(0, len(qb), ct)])
    got = np.frombuffer(c.recv_reply(want.size * 8 // ct, ct), dtype=np.uint64)
    assert c.at_eof()
    return got


def _client_driven(c, orc, crypto, d, alpha, dims, files, maxb):
    kind, cat, count, cmax = c.catalog()
    assert (kind, count, cmax) == (0, len(files), maxb)
    assert c.exchange_method() == CLIENT_DRIVEN
    assert c.optim_data() == (len(files), maxb)
    c.optim_exit()
    c.send_crypto_params(crypto + ":?")
    c.send_key(struct.pack(f"<{orc.m}Q", *orc.p))
    c.send_pir_params(d, alpha, dims)
    return [n for n, _ in cat]


@pytest.mark.parametrize("crypto,n,L,alpha,dims,no_pipeline", [
    ("LWE:80:2048:120", 30, 4096, 1, [6, 5], False),
    ("LWE:80:1024:60", 12, 3000, 3, [4], True),
    ("LWE:100:4096:180", 20, 2000, 1, [5, 4], False),  # three moduli, A > 56 absorb
    ("LWE:80:2048:120", 24, 1000, 1, [3, 4, 2], False),  # three levels
])
def test_client_driven_session(tmp_path, crypto, n, L, alpha, dims, no_pipeline):
    orc = Oracle(crypto, dims[0])
    db, files = _make_db(tmp_path, orc, n, L)
    extra = ["--no-pipeline"] if no_pipeline else []
    with Server(tmp_path, "--db-dir", db, "--sessions", 1, *extra) as srv:
        c = srv.client(timeout=120)
        c.hello()
        names = _client_driven(c, orc, crypto, len(dims), alpha, dims, files, L)
        index = len(names) // alpha - 1
        print("session:", crypto, dims, "alpha", alpha, "index", index, flush=True)
        sk = orc.keygen(4)
        q, want, elem, unit = _expected(orc, files, names, L, alpha, dims, sk, index, 3)
        got = _exchange_query(c, orc, q, want)
        c.close()
        assert srv.wait(120) == 0, srv.log()
    assert np.array_equal(got, want)
    assert extract_reply(orc, sk, got, dims, unit) == elem


def test_driven_mode_configures_then_serves_from_the_kept_database(tmp_path):
    """-z: the first client sets the parameters (saved to the file); the second is told
    them and is answered from the database the first session imported."""
    crypto, n, L, dims = "LWE:80:2048:120", 30, 4096, [6, 5]
    orc = Oracle(crypto, dims[0])
    db, files = _make_db(tmp_path, orc, n, L)
    cfg = tmp_path / "exp" / "PIRParams.cfg"
    cfg.parent.mkdir()
    with Server(tmp_path, "-z", cfg, "--db-dir", db, "--sessions", 2) as srv:
        sk = orc.keygen(4)
        c = srv.client(timeout=120)
        c.hello()
        names = _client_driven(c, orc, crypto, 2, 1, dims, files, L)
        print("session 1: configuring, index 7", flush=True)
        q, want, elem, unit = _expected(orc, files, names, L, 1, dims, sk, 7, 3)
        got = _exchange_query(c, orc, q, want)
        c.close()
        assert np.array_equal(got, want)
        assert extract_reply(orc, sk, got, dims, unit) == elem

        c = srv.client(timeout=120)
        c.hello()
        kind, cat, count, _ = c.catalog()
        assert [nm for nm, _ in cat] == names
        assert c.exchange_method() == SERVER_DRIVEN
        A = Oracle(crypto, dims[0]).A
        assert c.recv_crypto_params() == f"{crypto}:{A}"
        c.send_key(struct.pack(f"<{orc.m}Q", *orc.p))
        assert c.recv_pir_params() == (2, 1, dims)
        print("session 2: preimported database, index 22", flush=True)
        q, want, elem, unit = _expected(orc, files, names, L, 1, dims, sk, 22, 5)
        got = _exchange_query(c, orc, q, want)
        c.close()
        assert srv.wait(120) == 0, srv.log()
    assert np.array_equal(got, want)
    assert extract_reply(orc, sk, got, dims, unit) == elem
    assert cfg.read_text() == f"d 2\ndim 6 5 \n{crypto}:{A}\n"


def test_server_driven_from_params_file(tmp_path):
    crypto, n, L, dims = "LWE:80:2048:120", 20, 1500, [20]
    orc = Oracle(crypto, dims[0])
    db, files = _make_db(tmp_path, orc, n, L)
    cfg = tmp_path / "PIRParams.cfg"
    cfg.write_text(f"d 1\ndim 20 \n{crypto}\n")
    with Server(tmp_path, "-L", cfg, "--db-dir", db, "--sessions", 1) as srv:
        c = srv.client(timeout=120)
        c.hello()
        _, cat, _, _ = c.catalog()
        names = [nm for nm, _ in cat]
        assert c.exchange_method() == SERVER_DRIVEN
        assert c.recv_crypto_params() == crypto
        c.send_key(struct.pack(f"<{orc.m}Q", *orc.p))
        assert c.recv_pir_params() == (1, 1, dims)
        print("session: -L, index 13", flush=True)
        sk = orc.keygen(6)
        q, want, elem, unit = _expected(orc, files, names, L, 1, dims, sk, 13, 3)
        got = _exchange_query(c, orc, q, want)
        c.close()
        assert srv.wait(120) == 0, srv.log()
    assert np.array_equal(got, want)
    assert extract_reply(orc, sk, got, dims, unit) == elem


def test_concurrent_client_sessions(tmp_path):
    """three clients at once (--concurrent 3): uploads overlap, every session imports the
    database through the shared, locked handler, and each reply decrypts to its element"""
    import threading
    crypto, n, L, dims = "LWE:80:2048:120", 30, 4096, [6, 5]
    orc = Oracle(crypto, dims[0])
    db, files = _make_db(tmp_path, orc, n, L)
    sk = orc.keygen(4)
    indices = [3, 17, 29]
    results, errors = {}, []
    with Server(tmp_path, "--db-dir", db, "--sessions", 3, "--concurrent", 3) as srv:
        # the catalog order, from a first handshake, fixes what each index means
        clients = [srv.client(timeout=180) for _ in indices]
        names = None
        for c in clients:
            c.hello()
            names = _client_driven(c, orc, crypto, 2, 1, dims, files, L)
        expected = {i: _expected(orc, files, names, L, 1, dims, sk, i, 3 + i) for i in indices}

        def run(c, i):
            try:
                q, want, elem, unit = expected[i]
                results[i] = _exchange_query(c, orc, q, want)
            except Exception as e:  # reported below
                errors.append((i, repr(e)))

        print("sessions:", indices, flush=True)
        threads = [threading.Thread(target=run, args=(c, i)) for c, i in zip(clients, indices)]
        for t in threads:
            t.start()
        for t in threads:
            t.join(timeout=300)
        for c in clients:
            c.close()
        assert srv.wait(120) == 0, srv.log()
        assert srv.log().count("session done") == 3, srv.log()
    assert not errors, errors
    for i in indices:
        q, want, elem, unit = expected[i]
        assert np.array_equal(results[i], want), i
        assert extract_reply(orc, sk, results[i], dims, unit) == elem, i


def test_concurrent_sessions_on_the_kept_database(tmp_path):
    """-z with --concurrent 2: after the configuring client, two clients at once are
    answered from the one kept database (each session on its own pooled GPU context)"""
    import threading
    crypto, n, L, dims = "LWE:80:2048:120", 30, 4096, [6, 5]
    orc = Oracle(crypto, dims[0])
    db, files = _make_db(tmp_path, orc, n, L)
    cfg = tmp_path / "PIRParams.cfg"
    sk = orc.keygen(4)
    with Server(tmp_path, "-z", cfg, "--db-dir", db, "--sessions", 3, "--concurrent", 2) as srv:
        c = srv.client(timeout=180)
        c.hello()
        names = _client_driven(c, orc, crypto, 2, 1, dims, files, L)
        q, want, elem, unit = _expected(orc, files, names, L, 1, dims, sk, 5, 3)
        assert np.array_equal(_exchange_query(c, orc, q, want), want)
        c.close()
        indices = [11, 26]
        expected = {i: _expected(orc, files, names, L, 1, dims, sk, i, 7 + i) for i in indices}
        clients = []
        for _ in indices:
            c = srv.client(timeout=180)
            c.hello()
            c.catalog()
            assert c.exchange_method() == SERVER_DRIVEN
            c.recv_crypto_params()
            c.send_key(struct.pack(f"<{orc.m}Q", *orc.p))
            assert c.recv_pir_params() == (2, 1, dims)
            clients.append(c)
        results, errors = {}, []

        def run(c, i):
            try:
                q, want, _, _ = expected[i]
                results[i] = _exchange_query(c, orc, q, want)
            except Exception as e:  # reported below
                errors.append((i, repr(e)))

        print("sessions on the kept database:", indices, flush=True)
        threads = [threading.Thread(target=run, args=(c, i)) for c, i in zip(clients, indices)]
        for t in threads:
            t.start()
        for t in threads:
            t.join(timeout=300)
        for c in clients:
            c.close()
        assert srv.wait(120) == 0, srv.log()
    assert not errors, errors
    for i in indices:
        q, want, elem, unit = expected[i]
        assert np.array_equal(results[i], want), i
        assert extract_reply(orc, sk, results[i], dims, unit) == elem, i


def test_database_file_across_server_runs(tmp_path):
    """--db-file: the first server run imports the directory and writes the database file;
    a second run loads it instead of importing (same reply bytes); after one element file
    changes, a third run imports again (the file names another database) and rewrites it."""
    crypto, n, L, dims = "LWE:80:2048:120", 30, 4096, [6, 5]
    orc = Oracle(crypto, dims[0])
    db, files = _make_db(tmp_path, orc, n, L)
    ntt = tmp_path / "db.ntt"
    sk = orc.keygen(4)

    def run(run_no, index, seed):
        where = tmp_path / f"run{run_no}"  # the server's socket, log and working directory
        where.mkdir()
        with Server(where, "--db-dir", db, "--sessions", 1, "--no-pipeline", "--db-file", ntt) as srv:
            c = srv.client(timeout=120)
            c.hello()
            names = _client_driven(c, orc, crypto, 2, 1, dims, files, L)
            print(f"run {run_no}: index {index}", flush=True)
            q, want, elem, unit = _expected(orc, files, names, L, 1, dims, sk, index, seed)
            got = _exchange_query(c, orc, q, want)
            c.close()
            assert srv.wait(120) == 0, srv.log()
            log = srv.log()
        assert np.array_equal(got, want)
        assert extract_reply(orc, sk, got, dims, unit) == elem
        return log

    log1 = run(1, 7, 3)
    assert "loaded from the database file" not in log1 and ntt.exists()
    size1, mtime1 = ntt.stat().st_size, ntt.stat().st_mtime_ns
    log2 = run(2, 22, 5)
    assert "loaded from the database file" in log2, log2
    assert ntt.stat().st_mtime_ns == mtime1  # read, not rewritten
    # one element changes (same name and size, new bytes): imported again, file rewritten
    name = sorted(files)[3]
    newb = bytes((b + 1) & 0xff for b in files[name])
    (db / name).write_bytes(newb)
    files[name] = newb
    log3 = run(3, 3, 7)
    assert "loaded from the database file" not in log3, log3
    assert ntt.stat().st_size == size1 and ntt.stat().st_mtime_ns != mtime1
