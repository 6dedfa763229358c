"""The C-ABI library loads on the CPU, exports every symbol its header
declares, and the ctypes layouts match the C structs."""
import ctypes as C
import os
import re

from elegant_amd import _abi as A
from elegant_amd import native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "elegant_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*\b(eh_[a-z0-9_]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_header_symbol():
    lib = native.lib()
    names = header_functions()
    assert len(names) >= 19
    for n in names:
        assert hasattr(lib, n), n
    bound = {s[0] for s in native.SIGNATURES}
    assert set(names) == bound


def test_struct_layouts_match():
    lib = native.lib()
    assert lib.eh_struct_size(0) == C.sizeof(A.eh_element)
    for t, (_, cls) in A.PARAM_STRUCT.items():
        assert lib.eh_struct_size(t) == C.sizeof(cls), A.TYPE_NAMES[t]


def test_python_defaults_match_library_defaults():
    from elegant_amd import elements as E
    lib = native.lib()
    for t in A.PARAM_STRUCT:
        if t == A.EH_T_WAKE:
            continue
        ref = A.eh_element()
        assert lib.eh_element_defaults(C.byref(ref), t) == 0
        mine = E.make(t)
        field = A.PARAM_STRUCT[t][0]
        assert bytes(getattr(mine.p, field)) == bytes(getattr(ref.p, field)), A.TYPE_NAMES[t]


def test_strerror_and_no_device_error():
    lib = native.lib()
    assert lib.eh_strerror(-2).decode().startswith("element feature")
    import torch
    if not torch.cuda.is_available():
        ctx = C.c_void_p()
        rc = lib.eh_ctx_create(0, C.byref(ctx))
        assert rc == A.EH_ERR_NODEVICE
