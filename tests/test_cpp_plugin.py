"""The C++ boundary: the MI355X kernel plugin loaded through the host core's KernelProvider.

CPU tests check the plugin's C-linkage entry points, the provider's loading and error paths, and the
host core's graph operations that BreakBonds uses (edge removal, component split) and its reaction
descriptor parser on the reference's cases (TestContext.cpp:222-345);
the GPU test runs tests/cpp/test_plugin, which drives the C0 system and a bead-chain topology
through Context / StateModel / ActionFactory / SimulationLoop and checks them against the FP64
oracle (forces, energies, five UncontrolledApproximation steps event for event, conservation
under the device-fused Gillespie loop).
"""
import ctypes
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PLUGIN_DIR = os.path.join(ROOT, "readdy_amd", "plugins")
PLUGIN = os.path.join(PLUGIN_DIR, "libreaddy_amd_plugin.so")
TEST_BIN = os.path.join(ROOT, "tests", "cpp", "test_plugin")


def _need_build():
    if not (os.path.exists(PLUGIN) and os.path.exists(TEST_BIN)):
        pytest.skip("C++ host core not built (python -c 'import __graft_entry__ as g; g.build()')")


def test_plugin_exports_name_and_create_kernel():
    _need_build()
    lib = ctypes.CDLL(PLUGIN)
    lib.name.restype = ctypes.c_char_p
    assert lib.name() == b"MI355X"
    assert hasattr(lib, "createKernel")


def test_provider_loads_plugin_and_reports_errors():
    _need_build()
    out = subprocess.run([TEST_BIN, PLUGIN_DIR, "--provider-only"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "ALL PASSED" in out.stdout


def test_provider_rejects_missing_directory():
    _need_build()
    out = subprocess.run([TEST_BIN, os.path.join(ROOT, "no_such_dir"), "--provider-only"], capture_output=True,
                         text=True, timeout=120)
    assert out.returncode != 0
    assert "no such kernel plugin directory" in out.stdout


@pytest.mark.gpu
def test_cpp_plugin_against_oracle():
    _need_build()
    out = subprocess.run([TEST_BIN, PLUGIN_DIR], capture_output=True, text=True, timeout=300)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "ALL PASSED" in out.stdout


def test_host_core_graph_operations():
    """Edge removal and the connected-component split of the host core, on the CPU."""
    _need_build()
    out = subprocess.run([TEST_BIN, "--host-only"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "ALL PASSED" in out.stdout
