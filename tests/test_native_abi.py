"""C ABI checks that need no GPU: the library loads, exports every entry point the header
declares, and the ctypes structures match the C layout."""
from __future__ import annotations

import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "es_sieve.h")
HEADERS = sorted(os.path.join(ROOT, "include", f) for f in os.listdir(os.path.join(ROOT, "include"))
                 if f.endswith(".h"))


def header_functions(paths=HEADERS):
    names = set()
    for path in paths:
        text = open(path).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(es_[a-z_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    import es_native as nat
    L = nat.lib()
    names = header_functions()
    assert len(names) >= 14
    for n in names:
        assert hasattr(L, n), f"{n} declared in include/ but not exported"
    out = subprocess.run(["nm", "-D", "--defined-only", nat.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (es_[a-z_]+)\b", out))
    assert set(names) <= exported


def test_struct_layout_matches_c():
    import es_native as nat
    src = f"""
#include <stdio.h>
#include <stddef.h>
#include "{HEADER}"
int main(void) {{
  printf("%zu %zu %zu %zu %zu\\n", sizeof(es_plan_desc), offsetof(es_plan_desc, out_cap),
         offsetof(es_plan_desc, lds_budget), sizeof(es_engine_info),
         offsetof(es_engine_info, threads));
  return 0;
}}
"""
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        exe = os.path.join(d, "t")
        open(c, "w").write(src)
        subprocess.run(["gcc", c, "-o", exe], check=True)
        got = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    want = [C.sizeof(nat.PlanDesc), nat.PlanDesc.out_cap.offset, nat.PlanDesc.lds_budget.offset,
            C.sizeof(nat.EngineInfo), nat.EngineInfo.threads.offset]
    assert [int(x) for x in got] == want


def test_errors_without_device_are_reported():
    """With no GPU the ABI reports a device error instead of crashing."""
    import es_native as nat
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    n = C.c_int(-1)
    rc = nat.lib().es_device_query(C.byref(n), None, None, None, 0)
    assert rc == nat.ES_ERR_DEVICE
    assert nat.lib().es_last_error()


def test_create_rejects_bad_descriptors():
    """Validation happens before any device work, so it is testable on the CPU."""
    import es_native as nat
    L = nat.lib()
    d = nat.PlanDesc()
    d.abi_version = 99
    h = C.c_void_p()
    assert L.es_engine_create(C.byref(d), C.byref(h)) == nat.ES_ERR_INVALID
    assert b"ABI" in L.es_last_error()
    d.abi_version = nat.ES_ABI_VERSION
    d.n_rings = 0
    assert L.es_engine_create(C.byref(d), C.byref(h)) == nat.ES_ERR_INVALID
    assert L.es_engine_create(None, C.byref(h)) == nat.ES_ERR_INVALID


def test_engine_module_fails_loudly_without_library(tmp_path, monkeypatch):
    import es_native as nat
    monkeypatch.setattr(nat, "_lib", None)
    monkeypatch.setattr(nat, "LIB_PATH", str(tmp_path / "missing.so"))
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        nat.lib()
