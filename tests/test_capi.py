"""CPU checks of the C ABI: the library loads, exports every entry point that
include/xpir_amd.h declares, and its host-side parameter logic agrees with the oracle
and the reference's formulas.  No compute call needs a GPU here."""
import ctypes as C
import os
import re
import subprocess

import pytest

import xpir_amd
from xpir_amd import capi
from oracle.oracle import Oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "xpir_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(xpg_[a-z0-9_]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert declared_functions() == sorted(capi.EXPORTS)


def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", capi.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\bT (xpg_[a-z0-9_]+)", out))
    missing = set(declared_functions()) - exported
    assert not missing, missing
    lib = xpir_amd.lib()
    for name in declared_functions():
        assert hasattr(lib, name)


def test_struct_layouts_match_header(tmp_path):
    """The ctypes mirrors of the ABI's structs have the C compiler's sizes and offsets."""
    src = tmp_path / "layout.c"
    checks = {
        "xpg_tune_entry": (capi.XpgTuneEntry, ["n_l", "qn", "kernel", "c", "ms", "default_ms"]),
        "xpg_params": (capi.XpgParams, [f for f, _ in capi.XpgParams._fields_]),
    }
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', 'int main(void) {']
    for cname, (_, fields) in checks.items():
        lines.append(f'  printf("{cname} %zu\\n", sizeof({cname}));')
        for f in fields:
            lines.append(f'  printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines += ['  return 0;', '}']
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-o", str(exe), str(src)], check=True)
    got = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                       check=True).stdout.splitlines())
    for cname, (cls, fields) in checks.items():
        assert int(got[cname]) == C.sizeof(cls), cname
        for f in fields:
            assert int(got[f"{cname}.{f}"]) == getattr(cls, f).offset, f"{cname}.{f}"


@pytest.mark.parametrize("desc,n0", [("LWE:80:1024:60", 1000), ("LWE:80:2048:120", 65536),
                                     ("LWE:100:4096:180", 1024), ("LWE:80:2048:120", 7),
                                     ("LWE:128:8192:240", 100), ("LWE:80:512:60", 10)])
def test_params_match_oracle(desc, n0):
    prm = xpir_amd.params(desc, n0)
    orc = Oracle(desc, n0)
    assert (prm.n, prm.m, prm.A, prm.Berr) == (orc.n, orc.m, orc.A, orc.prm.Berr)
    assert [prm.moduli[i] for i in range(prm.m)] == orc.p
    assert prm.abs_bytes == orc.A * orc.n // 8
    assert prm.ct_bytes == 16 * orc.n * orc.m
    for nbytes in (1, 1000, 4096, 10240, 196608 * 3):
        assert xpir_amd.lib().xpg_polys_for_bytes(C.byref(prm), nbytes) == orc.polys(nbytes)
    P = 1
    for d in range(1, 4):
        assert xpir_amd.lib().xpg_reply_polys(C.byref(prm), 1, d) == P
        P = orc.next_polys(P)


@pytest.mark.parametrize("bad", ["", "LWE", "LWE:80:1000:60", "LWE:80:1024:70", "RSA:80:1024:60",
                                 "LWE:80:1024:300", "LWE:80:16384:60"])
def test_bad_parameter_strings_are_rejected(bad):
    with pytest.raises(capi.XpgError):
        xpir_amd.params(bad, 10)


def test_keygen_matches_oracle():
    prm = xpir_amd.params("LWE:100:4096:180", 1024)
    orc = Oracle("LWE:100:4096:180", 1024)
    assert (xpir_amd.keygen(prm, 77) == orc.keygen(77)).all()


def test_context_without_gpu_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(capi.XpgError):
        xpir_amd.Context(0)
