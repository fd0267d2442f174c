"""The session protocol of the server program (xpir_amd/_build/pir_server: the
reference's PIRSession and PIRServer, apps/server/) up to the query upload, driven by the
reference client's side of it (xpir_amd/pir_client.py): optimizer dry runs, catalogs, the
exchange of crypto and PIR parameters in both modes, the saved-parameters file, and
sessions that end early.  The query upload and the reply are in test_gpu_session.py."""
import os
import socket
import subprocess

import pytest

from xpir_amd.pir_client import CLIENT_DRIVEN, MEGA_BYTE, SERVER_DRIVEN, PIRClient

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SERVER = os.path.join(ROOT, "xpir_amd", "_build", "pir_server")


class Server:
    """pir_server accepting on a UNIX listening socket it inherits"""

    def __init__(self, tmp_path, *args):
        self.path = str(tmp_path / "pir.sock")
        self.lst = socket.socket(socket.AF_UNIX, socket.SOCK_STREAM)
        self.lst.bind(self.path)
        self.lst.listen(4)
        self.logf = open(tmp_path / "server.log", "w+")
        fd = self.lst.fileno()
        self.proc = subprocess.Popen([SERVER, "--listen-fd", str(fd), *map(str, args)], pass_fds=[fd],
                                     stdout=self.logf, stderr=subprocess.STDOUT, cwd=str(tmp_path))

    def client(self, timeout=30.0) -> PIRClient:
        return PIRClient(self.path, timeout)

    def wait(self, timeout=30.0) -> int:
        return self.proc.wait(timeout=timeout)

    def log(self) -> str:
        self.logf.flush()
        self.logf.seek(0)
        return self.logf.read()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        if self.proc.poll() is None:
            self.proc.kill()
            self.proc.wait()
        self.lst.close()
        self.logf.close()


def test_optimizer_dry_run(tmp_path):
    """is_client = 0: the optimizer's commands (DATA, ABS, SPEED, EXIT), then the next client."""
    exp = tmp_path / "exp"
    exp.mkdir()
    (exp / "preComputeLWE.abs").write_text("LWE:80:1024:60 1.3e-08 3.9e-08\n")
    with Server(tmp_path, "--db-generator", "-n", 7, "-l", 1234, "--optim-dir", exp) as srv:
        c = srv.client()
        c.hello(is_client=False)
        assert c.optim_data() == (7, 1234)
        assert c.optim_abs("LWE") == b"LWE:80:1024:60 1.3e-08 3.9e-08\n"
        assert c.optim_abs("Paillier") == b""  # no cache file: sent empty, as the reference does
        payload = os.urandom(MEGA_BYTE)
        assert c.optim_speed(payload) == payload
        c.optim_exit()
        assert c.at_eof()
        c.close()
        c = srv.client()  # dry runs do not end the server
        c.hello(is_client=False)
        assert c.optim_data() == (7, 1234)
        c.optim_exit()
        assert c.at_eof()
        c.close()


def test_client_driven_handshake(tmp_path):
    """catalog (full: names and sizes), CLIENT_DRIVEN, the optimizer round, the client's
    crypto parameters with key material, its PIR parameters; a client that then leaves
    ends the session without a reply."""
    db = tmp_path / "db"
    db.mkdir()
    sizes = {"a": 300, "b": 1000, "c": 20}
    for name, size in sizes.items():
        (db / name).write_bytes(os.urandom(size))
    with Server(tmp_path, "--db-dir", db, "--sessions", 1) as srv:
        c = srv.client()
        c.hello()
        kind, files, count, maxb = c.catalog()
        assert (kind, count, maxb) == (0, 3, 1000)
        assert dict(files) == sizes
        assert c.exchange_method() == CLIENT_DRIVEN
        assert c.optim_data() == (3, 1000)
        c.optim_exit()
        c.send_crypto_params("LWE:80:2048:120:?")
        c.send_key(bytes(16))
        c.send_pir_params(1, 1, [3])
        c.close()
        assert srv.wait() == 0
        assert "session failed" in srv.log()


def test_server_driven_from_params_file(tmp_path):
    """-L: the saved parameters are imposed on the client (SERVER_DRIVEN, crypto
    parameters, then d, alpha = 1 and the dimensions after the client's key)."""
    cfg = tmp_path / "PIRParams.cfg"
    cfg.write_text("d 2\ndim 4 3 \nLWE:80:2048:120:42\n")
    with Server(tmp_path, "-L", cfg, "--db-generator", "-n", 12, "-l", 3000, "--sessions", 1) as srv:
        c = srv.client()
        c.hello()
        kind, files, count, maxb = c.catalog()
        assert (kind, count, maxb) == (0, 12, 3000)
        assert [n for n, _ in files] == [str(i) for i in range(12)]
        assert c.exchange_method() == SERVER_DRIVEN
        assert c.recv_crypto_params() == "LWE:80:2048:120:42"
        c.send_key(bytes(16))
        assert c.recv_pir_params() == (2, 1, [4, 3])
        c.close()
        assert srv.wait() == 0


def test_compact_catalog_above_1000_elements(tmp_path):
    with Server(tmp_path, "--db-generator", "-n", 1001, "-l", 10, "--sessions", 1) as srv:
        c = srv.client()
        c.hello()
        assert c.catalog() == (1, [], 1001, 10)
        c.close()
        assert srv.wait() == 0


def test_unsupported_crypto_parameters_end_the_session(tmp_path):
    with Server(tmp_path, "--db-generator", "-n", 4, "-l", 100, "--sessions", 1) as srv:
        c = srv.client()
        c.hello()
        c.catalog()
        assert c.exchange_method() == CLIENT_DRIVEN
        c.optim_exit()
        c.send_crypto_params("LWE:80:1000:120")
        try:
            c.send_key(bytes(16))
        except (BrokenPipeError, ConnectionResetError):
            pass
        assert c.at_eof()
        c.close()
        assert srv.wait() == 0
        assert 'unsupported crypto parameters "LWE:80:1000:120"' in srv.log()


def test_grid_smaller_than_the_database_ends_the_session(tmp_path):
    """12 elements on a 3 x 3 grid: refused right after the PIR parameters, before any
    query byte (and before the GPU is touched)."""
    with Server(tmp_path, "--db-generator", "-n", 12, "-l", 100, "--sessions", 1) as srv:
        c = srv.client()
        c.hello()
        c.catalog()
        c.exchange_method()
        c.optim_exit()
        c.send_crypto_params("LWE:80:2048:120:?")
        c.send_key(bytes(16))
        c.send_pir_params(2, 1, [3, 3])
        assert c.at_eof()
        c.close()
        assert srv.wait() == 0
        assert "the PIR grid holds 9 elements, the database 12" in srv.log()


@pytest.mark.parametrize("content", [None, "d 3\ndim 4 3 \nLWE:80:2048:120\n", "d 1\ndim 4 \n"])
def test_bad_params_file(tmp_path, content):
    cfg = tmp_path / "PIRParams.cfg"
    if content is not None:
        cfg.write_text(content)
    r = subprocess.run([SERVER, "--fd", "0", "-L", str(cfg), "--db-generator"], capture_output=True, text=True,
                       timeout=30, stdin=subprocess.DEVNULL)
    assert r.returncode == 1
    assert "PIR parameters configuration file not found" in r.stderr


def test_usage_errors():
    for args in ([], ["--listen-fd", "3", "--fd", "0"], ["--fd", "0", "-p", "1234"], ["--fd", "x"]):
        r = subprocess.run([SERVER, *args], capture_output=True, text=True, timeout=30, stdin=subprocess.DEVNULL)
        assert r.returncode == 2, args


def test_concurrent_sessions(tmp_path):
    """--concurrent 2 serves a second client while the first one's session is still open;
    the default (one at a time) leaves the second waiting until the first ends."""
    import socket as _socket
    with Server(tmp_path, "--db-generator", "-n", 5, "-l", 64, "--concurrent", 2) as srv:
        a = srv.client()
        a.hello(is_client=False)
        assert a.optim_data() == (5, 64)  # a's session is open
        b = srv.client(timeout=10)
        b.hello(is_client=False)
        assert b.optim_data() == (5, 64)  # served beside it
        b.optim_exit()
        assert b.at_eof()
        a.optim_exit()
        assert a.at_eof()
        a.close()
        b.close()
    (tmp_path / "pir.sock").unlink()
    with Server(tmp_path, "--db-generator", "-n", 5, "-l", 64) as srv:
        a = srv.client()
        a.hello(is_client=False)
        assert a.optim_data() == (5, 64)
        b = srv.client(timeout=1.0)
        b.hello(is_client=False)
        b._pack("Q", 1)  # DATA: not answered while a's session holds the server
        with pytest.raises(_socket.timeout):
            b.recv(8)
        a.optim_exit()
        assert a.at_eof()
        b.s.settimeout(10)
        assert b.recv(16)  # now b's session runs
        b.optim_exit()
        a.close()
        b.close()
