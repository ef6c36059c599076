"""The C++ facade (include/tectonic.hpp) compiles against the C ABI, links
against libtectonic.so and reports a missing device as an exception instead
of exiting (the reference exits from CUDA_ERROR_CHECK)."""
import shutil
import subprocess

import pytest

from master_thesis_amd import _abi

SRC = r'''
#include <cstdio>
#include "tectonic.hpp"
int main()
{
    tt::PlateTectonicSim sim(512, 512, 1000u);
    if (sim.config().width != 512 || sim.config().num_plates != 16) return 2;
    try {
        sim.initialize(8, {});
    } catch (const tt::Error &e) {
        std::printf("error %d\n", e.status());
        return 0;
    }
    sim.executeIteration();
    std::printf("ran %u\n", sim.iteration());
    return 0;
}
'''


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_facade_builds_and_links(tmp_path, tectonic_lib):
    src = tmp_path / "host.cpp"
    src.write_text(SRC)
    exe = tmp_path / "host"
    root = _abi.PKG_DIR.parent
    subprocess.run(["g++", "-std=c++17", "-O1", f"-I{root / 'include'}", str(src), "-o", str(exe),
                    f"-L{_abi.PKG_DIR}", "-ltectonic", f"-Wl,-rpath,{_abi.PKG_DIR}"], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    # without a GPU the constructor works and initialize() throws a device error
    assert r.stdout.startswith("error -2") or r.stdout.startswith("ran 1")
