"""C ABI surface and host-side code, without a GPU: the library loads, exports
everything include/rtrex.h declares, the Escape loader follows the
reference's acceptance rules (GraphIO.cpp:19-74), and device entry points fail
loudly (status EDEVICE) when no gfx950 device is present."""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np
import pytest

import rtrex_cuda_amd as R
from conftest import ROOT


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "rtrex.h")).read()
    return sorted(set(re.findall(r"\b(rtrex_[a-z_0-9]+)\s*\(", txt)))


def test_exports_every_declared_symbol():
    L = C.CDLL(R.HIP_LIB)
    syms = header_symbols()
    assert set(syms) == set(R.EXPORTS)
    for s in syms:
        assert hasattr(L, s), s


def test_version():
    assert "gfx950" in R.version()


def test_footprint_grows():
    a = R.footprint_bytes(1000, 10000)
    b = R.footprint_bytes(2000, 20000)
    assert 0 < a < b


def _write(tmp_path, text):
    p = tmp_path / "g.txt"
    p.write_text(text)
    return str(p)


def test_escape_loader_rules(tmp_path):
    text = "# comment\n5 4\n\n0 1\n1 2 0.75\n# another\n2 3\n3 4 1.5 extra\n9 9\n"
    n, u, v = R.load_escape(_write(tmp_path, text))
    assert n == 5
    assert u.tolist() == [0, 1, 2, 3] and v.tolist() == [1, 2, 3, 4]


def test_escape_loader_truncated(tmp_path):
    with pytest.raises(R.RtrexError) as e:
        R.load_escape(_write(tmp_path, "4 3\n0 1\n1 2\n"))
    assert e.value.status == 4  # EIO, like ecIOError (GraphIO.cpp:67-71)


def test_escape_loader_missing_file(tmp_path):
    with pytest.raises(R.RtrexError) as e:
        R.load_escape(str(tmp_path / "absent.txt"))
    assert e.value.status == 4


def test_escape_loader_matches_fixture(tmp_path):
    import oracle as O
    from conftest import golden_graph
    n, u, v = golden_graph("small_test")
    p = str(tmp_path / "s.txt")
    O.write_escape(p, n, u, v)
    n2, u2, v2 = R.load_escape(p)
    assert n2 == n and np.array_equal(u2, u) and np.array_equal(v2, v)


def _gpu_present() -> bool:
    return os.path.exists("/dev/kfd") and any(
        os.path.exists(f"/dev/dri/renderD{128 + i}") for i in range(8))


@pytest.mark.skipif(_gpu_present(), reason="checks the no-GPU error path")
def test_device_calls_fail_loudly_without_gpu():
    with pytest.raises(R.RtrexError) as e:
        R.DeviceGraph.from_edges(3, np.array([0, 1], np.uint32), np.array([1, 2], np.uint32))
    assert e.value.status == 2
    assert "gfx950" in str(e.value) or "device" in str(e.value)


def test_cli_usage_and_no_device_error(tmp_path):
    """The command-line driver is built, prints its usage, and fails loudly
    (non-zero exit) where no gfx950 device is present."""
    import subprocess
    cli = os.path.join(os.path.dirname(R.HIP_LIB), "rtrex")
    r = subprocess.run([cli], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "Usage" in r.stdout
    if _gpu_present():
        return
    g = tmp_path / "g.txt"
    g.write_text("3 3\n0 1\n1 2\n0 2\n")
    r = subprocess.run([cli, str(g), "x", "0.3"], capture_output=True, text=True, timeout=60,
                       cwd=tmp_path)
    assert r.returncode == 1 and "gfx950" in r.stderr
