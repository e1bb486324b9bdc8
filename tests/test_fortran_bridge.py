"""The Fortran interface module compiles with the image's flang, and its bind(C) mirrors of
sh_model and shp_model have the sizes of the C structs (the ctypes mirrors are the C layout)."""
import os
import shutil
import subprocess

import pytest

from sunhubbardmy_amd import _lib
from sunhubbardmy_amd.proj import ShpModel

HERE = os.path.dirname(os.path.abspath(__file__))
BRIDGE = os.path.join(HERE, "..", "sunhubbardmy_amd", "fortran", "sunhip_bridge.f90")
FLANG = shutil.which("flang") or ("/opt/rocm/lib/llvm/bin/flang" if os.path.exists("/opt/rocm/lib/llvm/bin/flang") else None)

PROG = """program sizes
  use sunhip_bridge
  use, intrinsic :: iso_c_binding
  type(sh_model) :: a
  type(shp_model) :: b
  print '(i0, 1x, i0)', c_sizeof(a), c_sizeof(b)
end program sizes
"""


@pytest.mark.skipif(FLANG is None, reason="no flang in this image")
def test_bridge_compiles_and_struct_sizes_match(tmp_path):
    (tmp_path / "sizes.f90").write_text(PROG)
    subprocess.run([FLANG, "-c", os.path.abspath(BRIDGE), "-o", "bridge.o"], cwd=tmp_path, check=True,
                   capture_output=True)
    subprocess.run([FLANG, "sizes.f90", "bridge.o", "-o", "sizes"], cwd=tmp_path, check=True, capture_output=True)
    out = subprocess.run([str(tmp_path / "sizes")], capture_output=True, text=True, check=True).stdout.split()
    import ctypes
    assert int(out[0]) == ctypes.sizeof(_lib.ShModel)
    assert int(out[1]) == ctypes.sizeof(ShpModel)
