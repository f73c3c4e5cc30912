"""The batch driver's tile bookkeeping (csrc/oslmi_tiles.h), the host part
of the persistent multi-group launch and of its autotuner's trial loop,
built for the CPU with AddressSanitizer and UBSan and checked over random
batches and every tile size the autotuner tries (tests/hostbatch/
tiles_check.cpp)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "openshadinglanguage_amd", "csrc")


def test_tile_lists_under_sanitizers(tmp_path):
    exe = tmp_path / "tiles_check"
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined",
                           "-fno-sanitize-recover=all", "-I", CSRC, "-o", str(exe),
                           os.path.join(HERE, "hostbatch", "tiles_check.cpp")])
    # another library may be preloaded in the environment: ASan then runs
    # without insisting on being first in the link order
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0")
    out = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=120)
    assert out.returncode == 0, out.stderr[-2000:]
    assert out.stdout.startswith("ok ")
