"""The C-ABI library exports every entry point include/cubens.h declares, and the
Python mirror binds exactly those (no GPU calls here)."""
import ctypes as C
import os
import re

import pytest

import cubens

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "cubens.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(cb_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_the_abi():
    names = _declared()
    assert "cb_open" in names and "cb_run" in names and "cb_internal_force" in names
    assert set(names) == set(cubens.ABI), set(names) ^ set(cubens.ABI)


@pytest.mark.skipif(not os.path.exists(cubens.LIB_PATH), reason="HIP library not built")
def test_library_exports_every_symbol():
    lib = C.CDLL(cubens.LIB_PATH)
    for name in _declared():
        assert hasattr(lib, name), name


@pytest.mark.skipif(not os.path.exists(cubens.LIB_PATH), reason="HIP library not built")
def test_open_without_gpu_fails_loudly(tmp_path):
    """On a host without a GPU cb_open reports an error instead of computing on the CPU."""
    import torch  # noqa: F401  (device presence only)
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from cubens import inputs
    p = tmp_path / "m.txt"
    p.write_text(inputs.shell_panel(n=2).text)
    with pytest.raises(cubens.CubensError):
        cubens.GpuModel(str(p))


def test_bench_refuses_more_than_one_gpu():
    """bench.py is a one-GPU benchmark: any other --gpus exits 2 with a JSON error line."""
    import json
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], cwd=root,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]
