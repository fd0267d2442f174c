"""Optimizer cache lines measured on the GPU (xpir_amd/optim.py): one line per parameter
string in the reference's format, positive per-plaintext times, readable back."""
import pytest

pytestmark = pytest.mark.gpu


def test_abs_lines_on_the_gpu(gpu_ctx, tmp_path):
    from xpir_amd import optim
    params = ["LWE:80:1024:60", "LWE:80:2048:120"]
    path = tmp_path / "preComputeLWE.abs"
    lines = optim.write_abs(gpu_ctx, params, str(path))
    assert path.read_text() == "".join(x + "\n" for x in lines)
    assert optim.params_from_file(str(path)) == params
    for line, p in zip(lines, params):
        name, abs_s, pre_s = line.split(" ")
        assert name == p
        assert 0 < float(abs_s) < 1e-2 and 0 < float(pre_s) < 1e-2, line
