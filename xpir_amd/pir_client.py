"""The client side of the reference's session protocol (for tests, load generation and
Python callers of the server program), restated from
apps/client/PIRClient.cpp (hello :59-60, catalog :73-100, crypto parameters :215-266,
PIR parameters :137-163, exchange method :177, query upload :327-345, reply download
:387-420), apps/client/DESC.cpp:38-100 (catalog parsing) and the optimizer's client
commands (pir/PIROptimizerCommand.hpp:21, pir/replyGenerator/PIROptimizer.cpp:35-95,
226-284).  Integers are native little-endian: int/unsigned 4 bytes, short 2, size_t and
uint64 8."""
import socket
import struct

CLIENT_DRIVEN, SERVER_DRIVEN = 0, 1
ABS, DATA, SPEED, EXIT, NOP = range(5)
MEGA_BYTE = 1000000


class PIRClient:
    def __init__(self, path: str, timeout: float = 60.0):
        self.s = socket.socket(socket.AF_UNIX, socket.SOCK_STREAM)
        self.s.settimeout(timeout)
        self.s.connect(path)

    def close(self):
        self.s.close()

    # raw framing
    def send(self, b: bytes):
        self.s.sendall(b)

    def recv(self, n: int) -> bytes:
        out = bytearray()
        while len(out) < n:
            chunk = self.s.recv(min(n - len(out), 1 << 20))
            if not chunk:
                raise EOFError(f"stream ended after {len(out)} of {n} bytes")
            out += chunk
        return bytes(out)

    def at_eof(self) -> bool:
        """the server ended the stream (a reset counts: a server that closes with
        unread client bytes in its buffer resets the connection)"""
        try:
            return self.s.recv(1) == b""
        except ConnectionResetError:
            return True

    def _pack(self, fmt, *v):
        self.send(struct.pack("<" + fmt, *v))

    def _unpack(self, fmt):
        return struct.unpack("<" + fmt, self.recv(struct.calcsize("<" + fmt)))[0]

    # the session's messages, in protocol order
    def hello(self, is_client: bool = True):
        self._pack("i", 1 if is_client else 0)

    def catalog(self):
        """(type, [(name, size)] for the full catalog or [] for the compact one, count, max size)"""
        size = self._unpack("Q")
        text = self.recv(size).decode()
        lines = text.split("\n")
        kind, count = int(lines[0]), int(lines[1])
        if kind == 0:
            files = [(lines[2 + 2 * i], int(lines[3 + 2 * i])) for i in range(count)]
            return kind, files, count, max((s for _, s in files), default=0)
        return kind, [], count, int(lines[2])

    def exchange_method(self) -> int:
        return self._unpack("h")

    def optim_data(self):
        self._pack("Q", DATA)
        return self._unpack("Q"), self._unpack("Q")

    def optim_abs(self, crypto_name: str) -> bytes:
        self._pack("Q", ABS)
        self._pack("I", len(crypto_name))
        self.send(crypto_name.encode())
        return self.recv(self._unpack("i"))

    def optim_speed(self, payload: bytes) -> bytes:
        assert len(payload) == MEGA_BYTE
        self._pack("Q", SPEED)
        self.send(payload)
        return self.recv(MEGA_BYTE)

    def optim_exit(self):
        self._pack("Q", EXIT)

    def send_crypto_params(self, desc: str):
        self._pack("I", len(desc))
        self.send(desc.encode())

    def send_key(self, key: bytes):
        self._pack("I", len(key))
        self.send(key)

    def recv_crypto_params(self) -> str:
        return self.recv(self._unpack("i")).decode()

    def send_pir_params(self, d: int, alpha: int, dims):
        self._pack("i", d)
        self._pack("i", alpha)
        for n in dims:
            self._pack("i", n)

    def recv_pir_params(self):
        d = self._unpack("i")
        alpha = self._unpack("i")
        return d, alpha, [self._unpack("i") for _ in range(d)]

    def send_query(self, elements):
        for e in elements:
            self.send(e)

    def recv_reply(self, count: int, ct_bytes: int) -> bytes:
        return self.recv(count * ct_bytes)
