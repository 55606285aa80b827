"""The C ABI without a GPU: the library loads, exports every function that
include/psx_scft.h declares, the Python binding covers exactly that set, the
header is plain C, and the ctypes struct layouts match the C compiler's."""
import ctypes
import os
import re
import subprocess
import tempfile

import pytest

from polyswift_plusplus_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "psx_scft.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[a-z_]+\s*\*?\s*(psx_[A-Za-z0-9_]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_functions():
    names = declared_functions()
    assert "psx_fft_scaled_pair" in names and "psx_scft_iterate" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_abi.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header_exactly():
    assert sorted(_abi.EXPORTS) == declared_functions()


def _compile_and_run(src: str) -> str:
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "probe.c")
        exe = os.path.join(d, "probe")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                        "-o", exe, c], check=True, capture_output=True)
        return subprocess.run([exe], check=True, capture_output=True, text=True).stdout


def test_header_is_plain_c_and_layouts_match():
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "psx_scft.h"
int main(void) {
  printf("%zu %zu %zu\n", sizeof(psx_grid), offsetof(psx_grid, dx), offsetof(psx_grid, dz));
  printf("%zu %zu %zu\n", sizeof(psx_diblock), offsetof(psx_diblock, bB), offsetof(psx_diblock, chiN));
  printf("%zu %zu %zu %zu\n", sizeof(psx_mixing), offsetof(psx_mixing, depth),
         offsetof(psx_mixing, beta), offsetof(psx_mixing, f_order));
  printf("%zu %zu %zu %zu %zu %d %d\n", sizeof(psx_block), offsetof(psx_block, lengthfrac),
         offsetof(psx_block, nhead), offsetof(psx_block, head), offsetof(psx_block, tail),
         PSX_MAX_BLOCKS, PSX_MAX_JOIN);
  return 0;
}
'''
    out = _compile_and_run(src).split()
    g, c, m, b = _abi.Grid, _abi.Diblock, _abi.Mixing, _abi.BlockDesc
    want = [ctypes.sizeof(g), g.dx.offset, g.dz.offset,
            ctypes.sizeof(c), c.bB.offset, c.chiN.offset,
            ctypes.sizeof(m), m.depth.offset, m.beta.offset, m.f_order.offset,
            ctypes.sizeof(b), b.lengthfrac.offset, b.nhead.offset, b.head.offset, b.tail.offset,
            _abi.PSX_MAX_BLOCKS, _abi.PSX_MAX_JOIN]
    assert [int(x) for x in out] == want


def test_status_codes_match_header():
    text = open(HEADER).read()
    codes = dict((name, int(v)) for name, v in re.findall(r"(PSX_[A-Z_]+)\s*=\s*(\d+)", text))
    for k, name in _abi.STATUS_NAMES.items():
        assert codes[name] == k


def test_missing_library_fails_loudly(tmp_path):
    saved = _abi._lib
    try:
        _abi._lib = None
        with pytest.raises(RuntimeError, match="no CPU fallback"):
            _abi.load(str(tmp_path / "nope.so"))
    finally:
        _abi._lib = saved
