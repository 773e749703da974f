"""The C++ drop-in layer on the CPU: request-server semantics (the reference's
server tests, test/server/Server_Test.cpp, plus the drain contract and
concurrent clients) and DataManager error behaviour
(server/DataManager.cpp:19-60)."""
import ctypes as C
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "_bin")


def _need(path):
    if not os.path.exists(path):
        pytest.fail(f"{path} is not built: run __graft_entry__.build()")


def test_server_semantics():
    exe = os.path.join(BIN, "server_semantics")
    _need(exe)
    r = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=300)
    print(r.stdout)
    assert r.returncode == 0 and r.stdout.strip().endswith("ALL OK"), r.stdout


def test_datamanager_errors():
    so = os.path.join(BIN, "libservice_driver.so")
    _need(so)
    L = C.CDLL(so)
    buf = C.create_string_buffer(1024)
    assert L.drv_datamanager_errors(buf, 1024) == 0
    lines = buf.value.decode().splitlines()
    assert lines == ["Invalid DataItem (nullptr).", "Invalid id for DataItem (id = 3).",
                     "Invalid id for DataItem (id = 0)."]


def test_service_library_exports():
    so = os.path.join(ROOT, "gpuattract_2.0_amd", "lib", "libattract_service.so")
    _need(so)
    out = subprocess.run(["nm", "-DC", "--defined-only", so], stdout=subprocess.PIPE,
                         text=True).stdout
    for sym in ("as::MI355EnergyService6D<float>::createItemProcessor()",
                "as::MI355EnergyService6DModes<float>::createItemProcessor()",
                "as::DataManager::add(std::shared_ptr<as::DataItem>)"):
        assert sym in out, sym
