"""CPU: the batch engine's planning logic against a runtime emulator.

tree.hip and batch.hip are compiled with g++ (AddressSanitizer and UBSan
on) against tests/batch_emu: a host stand-in for the HIP runtime calls
they make (device memory is host memory, captured graphs replay recorded
launches) and stand-ins for the kernel launchers whose planes count steps.
emu_driver runs random scenario mixes (nests with REL_SIZE 2-5, REL_TIME
1-7 and a few in the thousands, which are run without graphs) in random
chunks, adds scenarios while others run and removes finished ones, and
checks that every scenario ends with exactly its own number of root steps
(REL_TIME times that in its child), continuity before momentum, JNQ/JNZ in
the right sub-steps, tiles covering each layer once per launch, queue heads
cleared, forked streams joined back before a capture ends, a fused momentum +
next continuity launch only where that step's continuity is done, and no
device memory left after a reset.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
EMU = os.path.join(HERE, "batch_emu")
CSRC = os.path.join(ROOT, "gpu-comcot_amd", "csrc")


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    # a directory per test process: parallel workers must not overwrite
    # each other's executable while it runs
    exe = str(tmp_path_factory.mktemp("batch_emu") / "emu_driver")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined",
           "-fno-omit-frame-pointer", "-I", EMU, "-I", CSRC, "-x", "c++",
           os.path.join(CSRC, "tree.hip"), os.path.join(CSRC, "batch.hip"), "-x", "none",
           os.path.join(EMU, "emu_runtime.cpp"), os.path.join(EMU, "emu_kernels.cpp"),
           os.path.join(EMU, "emu_driver.cpp"), "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


@pytest.mark.parametrize("seed", [1, 2, 3, 4, 5, 6, 7, 8])
def test_random_mixes(driver, seed):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1")
    r = subprocess.run([driver, str(seed), "25"], capture_output=True, text=True, env=env,
                       timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "verified" in r.stdout


def test_single_layer_scenarios_run_beside_the_nested_chain(driver):
    """Single-layer scenarios' launches fork onto a second stream and join
    back within each plan, captured into the graphs and enqueued directly,
    and every scenario still ends on its own step count."""
    import re
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1")
    r = subprocess.run([driver, "5", "25"], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    m = re.search(r"(\d+) cross-stream waits", r.stdout)
    assert m and int(m.group(1)) > 0, r.stdout


@pytest.mark.parametrize("seed", [21, 22])
def test_large_layers(driver, seed):
    env = dict(os.environ, EMU_BIG="1", ASAN_OPTIONS="detect_leaks=1")
    r = subprocess.run([driver, str(seed), "4"], capture_output=True, text=True, env=env,
                       timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]


@pytest.mark.parametrize("seed", [31, 32, 33])
def test_tuned_two_step_choice(driver, seed):
    """The tuner decides per plan whether linear roots take two steps per
    launch (emulated timings vary, so both decisions occur)."""
    env = dict(os.environ, EMU_TUNE_TWO="1", ASAN_OPTIONS="detect_leaks=1")
    r = subprocess.run([driver, str(seed), "25"], capture_output=True, text=True, env=env,
                       timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "two-step 1 (shape -1)" in r.stdout


@pytest.mark.parametrize("fuse", [1, 0])
@pytest.mark.parametrize("seed", [51, 52])
def test_fused_halves(driver, seed, fuse):
    """Nested linear roots whose momentum launch also computes the next root
    step's continuity: each such launch finds the continuity of its own
    step done (prologue or the previous fused launch), scenarios added
    later start without it, and every scenario ends on its own step count;
    switched off, none is fused."""
    import re
    env = dict(os.environ, EMU_FUSE=str(fuse), ASAN_OPTIONS="detect_leaks=1")
    r = subprocess.run([driver, str(seed), "25"], capture_output=True, text=True, env=env,
                       timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    m = re.search(r"fused halves (\d) \((\d+) tiles\)", r.stdout)
    assert m and int(m.group(1)) == fuse, r.stdout
    assert (int(m.group(2)) > 0) == (fuse == 1), r.stdout
