"""Host-side check of the kernels' special-form arithmetic (xpir_amd/csrc/arith.hpp):
the division-free Shoup quotient and the special-form lazy Shoup product against the
textbook 128-bit forms, for every modulus of the reference (crypto/NFLParams.cpp), on
boundary words and random words (tests/cpp/arith_check.cpp, built with hipcc as host
code; nothing runs on a GPU)."""
import os
import subprocess

from oracle.oracle import Oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_special_form_arithmetic(tmp_path):
    L = Oracle.lib()
    moduli = [str(L.xo_modulus(i)) for i in range(17)]
    exe = tmp_path / "arith_check"
    subprocess.run(["/opt/rocm/bin/hipcc", "-O2", "-std=c++17", "-x", "hip", "--offload-arch=gfx950",
                    "-o", str(exe), os.path.join(ROOT, "tests", "cpp", "arith_check.cpp")], check=True,
                   capture_output=True)
    r = subprocess.run([str(exe), "200000", *moduli], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stderr + r.stdout


# ---- the inline-asm sequences of arith.hpp, executed instruction by instruction ----
#
# The GPU parity tests pin what the kernels compute; this pins, on the CPU, that each
# asm sequence as written in arith.hpp computes the textbook value for every word class
# (edges and random words, all 17 moduli), that its result does not depend on the
# scratch registers' previous contents, and that every read of vcc as a lane mask or
# carry comes at least one wait state after the VALU write of vcc.

import random  # noqa: E402
import re  # noqa: E402

ARITH = os.path.join(ROOT, "xpir_amd", "csrc", "arith.hpp")
M32, M64 = (1 << 32) - 1, (1 << 64) - 1


def _asm_lines(fn):
    src = open(ARITH).read()
    i = src.index(f"{fn}(uint64_t") + len(fn)
    j = re.search(r"(?<!\w)asm\(", src[i:]).start() + i  # the asm statement, not the name
    k = re.search(r"\n\s*:", src[j:]).start() + j  # the first operand list
    text = ""
    m = re.match(r"asm\((\w+)", src[j:])
    if m:  # the template starts with a macro: its string literals first
        d = src.index(f"#define {m.group(1)}")
        e = src.index("\n", d)
        while src[e - 1] == "\\":
            e = src.index("\n", e + 1)
        text = "".join(re.findall(r'"((?:[^"\\]|\\.)*)"', src[d:e]))
    text += "".join(re.findall(r'"((?:[^"\\]|\\.)*)"', src[j:k]))
    return [s.strip() for s in text.split("\\n\\t") if s.strip()]


class _Lane:
    """One lane of a wave: 32-bit VGPRs v0..v15, named operands, vcc as this lane's bit."""

    def __init__(self, operands, rng):
        self.v = [rng.getrandbits(32) for _ in range(16)]  # scratch starts as garbage
        self.ops = dict(operands)  # name -> [value, bits]
        self.vcc, self.vcc_written, self.clock = rng.getrandbits(1), -10, 0

    def _read(self, tok, bits):
        if tok.startswith("%["):
            val, _ = self.ops[tok[2:-1]]
            return val
        m = re.fullmatch(r"v\[(\d+):(\d+)\]", tok)
        if m:
            a, b = int(m.group(1)), int(m.group(2))
            assert b == a + 1 and a % 2 == 0, f"misaligned pair {tok}"
            return self.v[a] | (self.v[b] << 32)
        m = re.fullmatch(r"v(\d+)", tok)
        if m:
            return self.v[int(m.group(1))]
        if tok == "vcc":
            assert self.clock - self.vcc_written >= 2, "vcc read without a wait state after its write"
            return self.vcc
        return int(tok, 0) & ((1 << bits) - 1)

    def _write(self, tok, val, bits):
        val &= (1 << bits) - 1
        if tok.startswith("%["):
            self.ops[tok[2:-1]][0] = val
            return
        m = re.fullmatch(r"v\[(\d+):(\d+)\]", tok)
        if m:
            a = int(m.group(1))
            self.v[a], self.v[a + 1] = val & M32, val >> 32
            return
        self.v[int(tok[1:])] = val

    def _set_vcc(self, bit):
        self.vcc, self.vcc_written = bit, self.clock

    def run(self, lines):
        for line in lines:
            op, _, rest = line.partition(" ")
            a = [t.strip() for t in rest.split(",")] if rest else []
            self.clock += 1
            if op == "s_nop":
                self.clock += int(a[0])  # s_nop N: N + 1 wait states
            elif op == "v_mad_u64_u32":
                res = self._read(a[2], 32) * self._read(a[3], 32) + self._read(a[4], 64)
                self._write(a[0], res, 64)
                self._set_vcc(res >> 64)
            elif op == "v_mul_lo_u32":
                self._write(a[0], self._read(a[1], 32) * self._read(a[2], 32), 32)
            elif op == "v_mov_b32":
                self._write(a[0], self._read(a[1], 32), 32)
            elif op == "v_cndmask_b32_e64":
                self._write(a[0], self._read(a[2], 32) if self._read(a[3], 1) else self._read(a[1], 32), 32)
            elif op == "v_add_u32":
                self._write(a[0], self._read(a[1], 32) + self._read(a[2], 32), 32)
            elif op == "v_addc_co_u32":  # VOP2: dst = src0 + src1 + vcc, carry-out to vcc
                assert a[1] == "vcc" and a[4] == "vcc"
                res = self._read(a[2], 32) + self._read(a[3], 32) + self._read("vcc", 1)
                self._write(a[0], res, 32)
                self._set_vcc(res >> 32)
            elif op == "v_add3_u32":
                self._write(a[0], self._read(a[1], 32) + self._read(a[2], 32) + self._read(a[3], 32), 32)
            elif op == "v_lshl_add_u64":
                self._write(a[0], (self._read(a[1], 64) << int(a[2])) + self._read(a[3], 64), 64)
            elif op == "v_lshlrev_b64":
                self._write(a[0], self._read(a[2], 64) << int(a[1]), 64)
            elif op == "v_cmp_gt_u64_e64":
                assert a[0] == "vcc"
                self._set_vcc(int(self._read(a[1], 64) > self._read(a[2], 64)))
            elif op == "v_sub_co_u32_e64":
                s0, s1 = self._read(a[2], 32), self._read(a[3], 32)
                self._write(a[0], s0 - s1, 32)
                self._set_vcc(int(s0 < s1))
            elif op == "v_subb_co_u32_e64":
                s0, s1, b = self._read(a[2], 32), self._read(a[3], 32), self._read(a[4], 1)
                self._write(a[0], s0 - s1 - b, 32)
                self._set_vcc(int(s0 < s1 + b))
            else:
                raise AssertionError(f"instruction not modelled: {line}")
        return self


def _lazy(x, w, wq, p):
    return (x * w - ((x * wq) >> 64) * p) & M64


def _words(rng, p):
    x = rng.choice([rng.getrandbits(64), rng.randrange(2 * p), 0, 1, 2 * p - 1, M64, p])
    y = rng.choice([rng.randrange(p), 0, 1, p - 1])
    return x, y, ((y << 64) // p) & M64


def _split(x):
    return x & M32, x >> 32


def test_asm_sequences_compute_the_textbook_values():
    L = Oracle.lib()
    rng = random.Random(20261204)
    mac, lazy, add, sub = (_asm_lines(f) for f in ("mac_step_asm", "shoup_lazy_asm", "add_mod2p_asm", "sub_mod2p_asm"))
    mac_count = _asm_lines("mac_step_count_asm")
    # 17 and 12 VALU instructions and one s_nop each; with T's v_mul_hi_u32 from C, 18 and 13;
    # the counting step adds one add-with-carry
    assert sum(ln.startswith("v_") for ln in mac) == 17 and sum(ln.startswith("v_") for ln in lazy) == 12
    assert mac_count[:-1] == mac and mac_count[-1].startswith("v_addc_co_u32")
    for i in range(17):
        p = L.xo_modulus(i)
        d = (1 << 61) - p
        for _ in range(600):
            x, y, yq = _words(rng, p)
            xl, xh = _split(x)
            yl, yh = _split(y)
            ql, qh = _split(yq)
            T = (xl * ql) >> 32
            # chain step: acc + r, minus p when that reaches p
            acc = rng.choice([rng.randrange(2 * p), 0, p - 1, p, 2 * p - 1, rng.getrandbits(62)])
            ops = dict(acc=[acc, 64], xl=[xl, 32], xh=[xh, 32], yl=[yl, 32], yh=[yh, 32], ql=[ql, 32],
                       qh=[qh, 32], T=[T, 64], d=[d, 32], k29=[0xE0000000, 32], negp=[(-p) & M64, 64],
                       plo=[p & M32, 32], phi=[p >> 32, 32])
            got = _Lane(ops, rng).run(mac).ops["acc"][0]
            want = (acc + _lazy(x, y, yq, p)) & M64
            subtracted = want >= p
            want = want - p if want >= p else want
            assert got == want, (i, x, y, acc)
            # the counting form: the same word, and the count grows exactly when p was subtracted
            ns = rng.randrange(1000)
            lane = _Lane(dict(ops, acc=[acc, 64], ns=[ns, 32]), rng).run(mac_count)
            assert lane.ops["acc"][0] == want and lane.ops["ns"][0] == ns + subtracted, (i, x, y, acc)
            # lazy product
            ops = dict(r=[0, 64], xl=[xl, 32], xh=[xh, 32], wl=[yl, 32], wh=[yh, 32], ql=[ql, 32], qh=[qh, 32],
                       T=[T, 64], d=[d, 32], k29=[0xE0000000, 32])
            assert _Lane(ops, rng).run(lazy).ops["r"][0] == _lazy(x, y, yq, p), (i, x, y)
            # butterfly adds on lazy words in [0, 2p)
            u, v = rng.randrange(2 * p), rng.choice([rng.randrange(2 * p), 0, 2 * p - 1])
            ops = dict(lo=[0, 32], hi=[0, 32], u=[u, 64], v=[v, 64], m2=[2 * p, 64], negm2=[(-2 * p) & M64, 64])
            lane = _Lane(ops, rng).run(add)
            s = u + v
            assert lane.ops["lo"][0] | (lane.ops["hi"][0] << 32) == (s - 2 * p if s >= 2 * p else s)
            ul, uh = _split(u)
            vl, vh = _split(v)
            ops = dict(lo=[0, 32], hi=[0, 32], ul=[ul, 32], uh=[uh, 32], vl=[vl, 32], vh=[vh, 32], m2=[2 * p, 64])
            lane = _Lane(ops, rng).run(sub)
            assert lane.ops["lo"][0] | (lane.ops["hi"][0] << 32) == ((u - v) & M64 if u >= v else u - v + 2 * p)


def test_asm_model_catches_a_planted_mistake():
    """The interpreter is not vacuous: the chain step with its select operands swapped
    (keep u + p when the reference keeps u) disagrees with the textbook step."""
    L = Oracle.lib()
    rng = random.Random(7)
    mac = [ln.replace("v_cndmask_b32_e64 v2, %[plo], 0, vcc", "v_cndmask_b32_e64 v2, 0, %[plo], vcc")
           for ln in _asm_lines("mac_step_asm")]
    p = L.xo_modulus(0)
    d = (1 << 61) - p
    bad = 0
    for _ in range(200):
        x, y, yq = _words(rng, p)
        xl, xh = _split(x)
        yl, yh = _split(y)
        ql, qh = _split(yq)
        acc = rng.randrange(2 * p)
        ops = dict(acc=[acc, 64], xl=[xl, 32], xh=[xh, 32], yl=[yl, 32], yh=[yh, 32], ql=[ql, 32], qh=[qh, 32],
                   T=[(xl * ql) >> 32, 64], d=[d, 32], k29=[0xE0000000, 32], negp=[(-p) & M64, 64],
                   plo=[p & M32, 32], phi=[p >> 32, 32])
        want = (acc + _lazy(x, y, yq, p)) & M64
        want = want - p if want >= p else want
        bad += _Lane(ops, rng).run(mac).ops["acc"][0] != want
    assert bad > 0
