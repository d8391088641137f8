"""The C-ABI library loads and exports every symbol include/sgipc.h declares
(no device calls)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "sgipc.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(sgipc_[a-z_0-9]+)\s*\(", src)))


def test_library_exports_all_declared_symbols():
    lib = ctypes.CDLL(os.path.join(ROOT, "stiff-gipc_amd", "libsgipc.so"))
    names = _declared()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_lists_the_same_symbols():
    from sgipc import _lib
    assert sorted(_lib.EXPORTS) == _declared()


def test_scene_desc_layout_matches_header():
    """ctypes field order equals the C struct's member order"""
    from sgipc.scene import SceneDesc
    src = open(os.path.join(ROOT, "include", "sgipc_scene.h")).read()
    body = src[src.index("typedef struct sgipc_scene_desc"):src.index("} sgipc_scene_desc;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    names = re.findall(r"\b([A-Za-z_][A-Za-z_0-9]*)(?:\[3\])?\s*[;,]", body)
    fields = [f for f, _ in SceneDesc._fields_]
    assert names == fields


def test_last_error_follows_the_last_call(tmp_path):
    """a failing call leaves its message, a successful one clears it (host-only
    entry points, no device)"""
    import numpy as np
    import sgipc
    from sgipc import lib
    with __import__("pytest").raises(sgipc.SgipcError):
        sgipc.read_mm_vector(tmp_path / "missing.mtx")
    assert b"missing.mtx" in lib().sgipc_last_error()
    p = tmp_path / "b.mtx"
    p.write_text("%%MatrixMarket matrix array real general\n2 1\n1.5\n-2\n")
    assert np.array_equal(sgipc.read_mm_vector(p), [1.5, -2.0])
    assert lib().sgipc_last_error() == b""
