"""CPU checks of xpir_amd/optim.py: the reference optimizer's doubling search
(PIROptimizer.cpp:155-207) and its cache-line format (5 significant digits,
std::setprecision(5), PIROptimizer.cpp:142-145), compared with the C++ stream output."""
import subprocess

from xpir_amd import optim


def test_doubling_loop_stops_at_half_a_second():
    seen = []

    def sim(p):
        seen.append(p)
        return p * 1e-3

    per = optim.doubling_loop(sim, 4, 128)
    assert seen == [4, 8, 16, 32, 64, 128, 256, 512]
    assert per == 0.512 / (4 * 512)


def test_doubling_loop_stops_at_the_size_cap():
    seen = []
    ct_bits = 128 * 1024  # LWE:*:1024:60 ciphertext bits
    per = optim.doubling_loop(lambda p: seen.append(p) or 1e-6 * p, 4, ct_bits)
    # the last size run: p * 2 * 4 * ct_bits reaches 2^30
    assert seen[-1] == 1024 and per == 1e-6 * 1024 / (4 * 1024)


def test_precompute_loop_stops_on_zero():
    seen = []
    per = optim.doubling_loop(lambda p: seen.append(p) or 0.0, 128, 1 << 17, stop_on_zero=True)
    assert seen == [128] and per == 0.0


def test_line_format_matches_cpp_setprecision(tmp_path):
    vals = [2.5e-4, 1.23456789e-7, 0.5, 3.0, 12345678.0, 1.0e-12, 0.000123456]
    src = tmp_path / "fmt.cpp"
    src.write_text(
        "#include <iostream>\n#include <iomanip>\n#include <sstream>\nint main(){double v[]={"
        + ",".join(repr(v) for v in vals)
        + "};for(double x: v){std::ostringstream o; o<<std::setprecision(5)<<x<<\" \"<<x; std::cout<<o.str()<<\"\\n\";}}\n")
    exe = tmp_path / "fmt"
    subprocess.run(["g++", "-o", str(exe), str(src)], check=True)
    cpp = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines()
    for v, line in zip(vals, cpp):
        assert optim.format_line("LWE:80:1024:60", v, v) == "LWE:80:1024:60 " + line


def test_params_from_file(tmp_path):
    p = tmp_path / "preComputeLWE.abs"
    p.write_text("LWE:80:1024:60 1.2e-05 3e-06\nLWE:80:2048:120 2e-05 4e-06\n")
    assert optim.params_from_file(str(p)) == ["LWE:80:1024:60", "LWE:80:2048:120"]
