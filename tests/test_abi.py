"""C ABI of libjaccard_amd.so on the CPU: the library loads, exports every
function include/jaccard_amd.h declares, and rejects bad arguments before
touching a device (no compute calls here)."""
import ctypes
import os
import re

import pytest

from jaccard_amd import _native, api

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "jaccard_amd.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(jc_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_reference_entry_points():
    names = declared()
    for must in ("jc_jaccard", "jc_jaccard_list", "jc_rowsum", "jc_fill", "jc_intersect", "jc_intersect_pairs",
                 "jc_jw", "jc_ec_scan", "jc_ec_unweighted", "jc_vertex_weights", "jc_nonzero_count"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = _native.load_jaccard_lib()
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_one_hip_runtime_in_process():
    """Loading the library before anything imported torch still leaves ONE
    HIP runtime in the process (torch's): a second runtime would own none of
    torch's tensors or streams."""
    import subprocess
    import sys
    code = (
        "import sys; sys.path.insert(0, %r)\n"
        "from jaccard_amd import _native\n"
        "_native.load_jaccard_lib()\n"
        "import torch; torch.cuda.is_available()\n"
        "maps = {l.split()[-1] for l in open('/proc/self/maps') if 'libamdhip64' in l}\n"
        "print(len(maps)); print(*sorted(maps), sep='\\n')\n" % os.path.join(ROOT, "sycl-jaccard_amd"))
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, check=True).stdout
    assert out.splitlines()[0] == "1", out


def test_header_compiles_as_c():
    import subprocess
    import tempfile
    src = '#include "jaccard_amd.h"\nint main(void){ return JC_OK; }\n'
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "t.c")
        open(p, "w").write(src)
        subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.dirname(HEADER), "-c", p, "-o",
                        os.path.join(d, "t.o")], check=True)


def test_argument_validation_without_device():
    lib = _native.load_jaccard_lib()
    assert lib.jc_jaccard(None, None, None, 1, None) == api.ERR_INVALID_ARG
    bad = api.JcCsr(-1, 0, None, None, None)
    assert lib.jc_jaccard(ctypes.byref(bad), None, None, 1, None) == api.ERR_INVALID_ARG
    nullind = api.JcCsr(4, 3, 1234, None, None)
    assert lib.jc_jaccard(ctypes.byref(nullind), None, 5678, 1, None) == api.ERR_INVALID_ARG
    empty = api.JcCsr(0, 0, None, None, None)
    assert lib.jc_jaccard(ctypes.byref(empty), None, None, 1, None) == 0
    assert lib.jc_jaccard_ex(ctypes.byref(empty), None, None, 1, 8, None, 0, None) == api.ERR_INVALID_ARG
    assert lib.jc_jaccard_list(ctypes.byref(empty), None, -1, None, None, None, None) == api.ERR_INVALID_ARG
    assert lib.jc_fill(-1, None, 0.0, None) == api.ERR_INVALID_ARG
    assert lib.jc_jw(-2, None, None, None, None) == api.ERR_INVALID_ARG
    assert lib.jc_rowsum(3, 0, None, None, None, None, None) == api.ERR_INVALID_ARG
    assert lib.jc_nonzero_count(4, None, 1e-5, None, None) == api.ERR_INVALID_ARG


def test_strerror_and_version():
    lib = _native.load_jaccard_lib()
    for code in (0, -1, -2, -3, -4, -5, -99):
        assert lib.jc_strerror(code)
    assert b"gfx950" in lib.jc_version()


def test_workspace_size_grows_with_graph():
    a = api.workspace_size(1000, 10000, False, True)
    b = api.workspace_size(100000, 1000000, False, True)
    c = api.workspace_size(100000, 1000000, True, False)
    assert 0 < a < b < c


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_native, "_jaccard", None)
    monkeypatch.setattr(_native, "JACCARD_LIB", str(tmp_path / "absent.so"))
    with pytest.raises(_native.NativeLibraryMissing):
        _native.load_jaccard_lib()


def _host_graph(n=4, m=6):
    """A DeviceCSR shell over host tensors: the Python layer must refuse every
    bad argument before the library is called, so no device is needed."""
    import torch
    g = api.DeviceCSR.__new__(api.DeviceCSR)
    g.n, g.offsets, g.indices, g.edge_w = n, torch.zeros(n + 1, dtype=torch.int32), torch.zeros(m, dtype=torch.int32), None
    return g


@pytest.mark.parametrize("case", ["wrong_dtype", "too_short", "host_tensor", "not_tensor"])
def test_python_entry_points_check_tensors(case):
    import torch
    g = _host_graph()
    f32 = {"wrong_dtype": torch.zeros(6, dtype=torch.float64), "too_short": torch.zeros(3, dtype=torch.float32),
           "host_tensor": torch.zeros(6, dtype=torch.float32), "not_tensor": [0.0] * 6}[case]
    i32 = {"wrong_dtype": torch.zeros(3, dtype=torch.int64), "too_short": torch.zeros(3, dtype=torch.int32),
           "host_tensor": torch.zeros(3, dtype=torch.int32), "not_tensor": [0] * 3}[case]
    dest = {"wrong_dtype": torch.zeros(6, dtype=torch.int64), "too_short": torch.zeros(3, dtype=torch.int32),
            "host_tensor": torch.zeros(6, dtype=torch.int32), "not_tensor": [0] * 6}[case]
    msg = {"wrong_dtype": "must be torch", "too_short": "at least", "host_tensor": "device",
           "not_tensor": "torch tensor"}[case]
    calls = [
        lambda: api.jaccard(g, result=f32),
        lambda: api.vertex_weights(g, out=f32),
        lambda: api.rowsum(g, f32),
        lambda: api.intersect(g, f32, f32, f32),
        lambda: api.ec_scan(g, dest, f32),
        lambda: api.ec_unweighted(g, None, f32),
    ]
    if case != "too_short":  # these take their length from the first tensor itself
        calls += [lambda: api.nonzero_count(f32), lambda: api.fill(f32, 0.0), lambda: api.jw(f32, f32, f32),
                  lambda: api.jaccard_list(g, i32, i32), lambda: api.csr_from_edges(4, i32, i32)]
    for call in calls:
        with pytest.raises(ValueError, match=msg):
            call()


def test_python_weights_checked():
    import torch
    g = _host_graph()
    with pytest.raises(ValueError, match="weights needs at least 4"):
        api.jaccard(g, result=torch.zeros(6, dtype=torch.float32), weights=torch.zeros(3))
    with pytest.raises(ValueError, match="same length"):
        api.jaccard_list(g, torch.zeros(3, dtype=torch.int32), torch.zeros(2, dtype=torch.int32))
