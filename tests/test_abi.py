"""The C ABI library loads, exports every symbol include/gtls_amd.h declares, and the ctypes
mirror of struct gtls_batch has the C layout. No device calls (runs without a GPU)."""
import ctypes as ct
import os
import re
import subprocess

import pytest

from gtls_amd import _ffi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gtls_amd.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(gtls_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_ffi.lib_path()):
        pytest.skip("libgtls_amd.so not built")
    lib = _ffi.load()
    names = declared_functions()
    assert names, "no declarations parsed"
    assert sorted(names) == sorted(_ffi.EXPORTED)
    for name in names:
        assert hasattr(lib, name), name


def test_ctypes_struct_matches_c_layout(tmp_path):
    fields = [f[0] for f in _ffi.GtlsBatch._fields_]
    prog = tmp_path / "layout.c"
    body = "\n".join('printf("%s %%zu\\n", offsetof(gtls_batch, %s));' % (f, f) for f in fields)
    prog.write_text('#include <stddef.h>\n#include <stdio.h>\n#include "gtls_amd.h"\n'
                    'int main(void){%s printf("sizeof %%zu\\n", sizeof(gtls_batch));return 0;}'
                    % body)
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I" + os.path.join(ROOT, "include"), str(prog), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    c_off = dict(line.split() for line in out.strip().splitlines())
    for f in fields:
        assert int(c_off[f]) == getattr(_ffi.GtlsBatch, f).offset, f
    assert int(c_off["sizeof"]) == ct.sizeof(_ffi.GtlsBatch)


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_ffi, "_LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(_ffi, "_lib", None)
    with pytest.raises(ImportError):
        _ffi.load()
