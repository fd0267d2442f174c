"""Host-only parts of the C++ facade (xpir_amd/cpp/libpir_amd.hpp), built and run without
a GPU: DBDirectoryProcessor (one file per element and file splitting), its catalog
strings and zero padding, DBHandler::readAggregatedStream, and the block reads the
import uses (readStreams) against element-by-element reads (tests/cpp/db_directory.cpp)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "xpir_amd")


def test_db_directory_processor(tmp_path):
    exe = tmp_path / "db_directory"
    subprocess.run(["g++", "-O1", "-std=c++17", "-Wall", "-pthread", "-o", str(exe),
                    os.path.join(ROOT, "tests", "cpp", "db_directory.cpp"),
                    os.path.join(PKG, "cpp", "libpir_amd.cpp"),
                    "-L", PKG, "-lxpir_amd", f"-Wl,-rpath,{PKG}"], check=True)
    r = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stderr + r.stdout
