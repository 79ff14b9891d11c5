"""The C ABI library loads without a GPU and exports every declared symbol."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "onionpir_gpu.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(opir_[a-z0-9_]+)\s*\(", src)))


def test_header_matches_binding_list():
    from onionpirv2_amd import gpu
    assert declared_functions() == sorted(gpu.EXPORTED)


def test_library_exports_every_symbol():
    from onionpirv2_amd import gpu
    lib = gpu.load_library()
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_ctx_create_rejects_bad_scheme_without_gpu():
    """Argument validation happens before any device call."""
    from onionpirv2_amd import gpu
    lib = gpu.load_library()
    s = gpu.OpirScheme(poly_degree=4096, composite=1)
    h = ctypes.c_void_p()
    assert lib.opir_ctx_create(ctypes.byref(s), 0, ctypes.byref(h)) == gpu.OPIR_EINVAL
    assert b"composite" in lib.opir_last_error(None)
    assert lib.opir_answer(None, 0, None, None) == gpu.OPIR_EINVAL
