"""The Fortran interface module compiles with the image's flang, and its bind(C) mirrors of
sh_model and shp_model have the sizes and the component offsets of the C structs (the ctypes
mirrors are the C layout)."""
import os
import shutil
import subprocess

import pytest

from sunhubbardmy_amd import _lib
from sunhubbardmy_amd.proj import ShpModel

HERE = os.path.dirname(os.path.abspath(__file__))
BRIDGE = os.path.join(HERE, "..", "sunhubbardmy_amd", "fortran", "sunhip_bridge.f90")
FLANG = shutil.which("flang") or ("/opt/rocm/lib/llvm/bin/flang" if os.path.exists("/opt/rocm/lib/llvm/bin/flang") else None)

def _offsets_program():
    """A Fortran program printing the size of both structs, then the byte offset of every
    component (c_loc of the component minus c_loc of the struct), in the ctypes field order."""
    lines = ["program sizes", "  use sunhip_bridge", "  use, intrinsic :: iso_c_binding",
             "  type(sh_model), target :: a", "  type(shp_model), target :: b",
             "  print '(i0, 1x, i0)', c_sizeof(a), c_sizeof(b)"]
    for var, cls in (("a", _lib.ShModel), ("b", ShpModel)):
        for name, _ in cls._fields_:
            lines.append("  print '(i0)', transfer(c_loc(%s%%%s), 0_c_intptr_t) - transfer(c_loc(%s), 0_c_intptr_t)"
                         % (var, name, var))
    lines.append("end program sizes")
    return "\n".join(lines) + "\n"


@pytest.mark.skipif(FLANG is None, reason="no flang in this image")
def test_bridge_compiles_and_struct_sizes_match(tmp_path):
    (tmp_path / "sizes.f90").write_text(_offsets_program())
    subprocess.run([FLANG, "-c", os.path.abspath(BRIDGE), "-o", "bridge.o"], cwd=tmp_path, check=True,
                   capture_output=True)
    subprocess.run([FLANG, "sizes.f90", "bridge.o", "-o", "sizes"], cwd=tmp_path, check=True, capture_output=True)
    out = subprocess.run([str(tmp_path / "sizes")], capture_output=True, text=True, check=True).stdout.split()
    import ctypes
    assert int(out[0]) == ctypes.sizeof(_lib.ShModel)
    assert int(out[1]) == ctypes.sizeof(ShpModel)
    # every component at the C offset: a swap of two members of one size moves two offsets
    expect = [getattr(cls, name).offset for cls in (_lib.ShModel, ShpModel) for name, _ in cls._fields_]
    assert [int(v) for v in out[2:]] == expect
