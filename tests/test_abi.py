"""The C-ABI library loads and exports every entry point include/*.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if fn.endswith(".h"):
            text = open(os.path.join(inc, fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(pm_[a-z_0-9]+)\s*\(", text))
    return names


def _lib():
    from protomol_amd import _lib as L
    if not os.path.exists(L.LIB_PATH):
        from protomol_amd import build
        build.build()
    return ctypes.CDLL(L.LIB_PATH)


def test_header_declares_entry_points():
    names = _declared()
    assert {"pm_ctx_create", "pm_evaluate", "pm_md_run", "pm_last_error"} <= names


def test_library_exports_every_declared_symbol():
    lib = _lib()
    missing = [n for n in sorted(_declared()) if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_names_match_header():
    from protomol_amd import _lib as L
    assert set(L.EXPORTED) == _declared()


def test_abi_version_without_gpu():
    lib = _lib()
    text = open(os.path.join(ROOT, "include", "protomol_amd.h")).read()
    assert lib.pm_abi_version() == int(re.search(r"#define PM_ABI_VERSION (\d+)", text).group(1))


def test_struct_layouts_match_header():
    """ctypes mirrors carry the header's field order and sizes."""
    from protomol_amd import _lib as L
    text = open(os.path.join(ROOT, "include", "protomol_amd.h")).read()
    for struct, cls in (("pm_switch_desc", L.SwitchDesc), ("pm_force_desc", L.ForceDesc),
                        ("pm_system_desc", L.SystemDesc)):
        body = re.search(r"typedef struct %s \{(.*?)\} %s;" % (struct, struct), text, re.S).group(1)
        body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
        fields = []
        for decl in body.split(";"):
            decl = decl.strip()
            if not decl:
                continue
            names = re.findall(r"\*?\s*([a-z_0-9]+)\s*(?:\[\d+\])?\s*(?:,|$)", decl)
            fields += names
        assert [f[0] for f in cls._fields_] == fields, struct
