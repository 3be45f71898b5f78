"""The iso_c_binding module a Tonto build would use compiles with a Fortran compiler and
links against libtonto_gpu.so (no GPU calls)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FC = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"


@pytest.mark.skipif(not os.path.exists(FC), reason="no Fortran compiler in this image")
def test_fortran_module_compiles_and_links(tmp_path):
    src = os.path.join(ROOT, "tonto_amd", "fortran", "tonto_gpu_mod.f90")
    prog = tmp_path / "use_it.f90"
    prog.write_text("program use_it\n  use tonto_gpu_mod\n  type(c_ptr) :: ctx\n"
                    "  real(c_double) :: P(2,2), J(2,2), K(2,2)\n  P = 0\n"
                    "  if (.false.) call gpu_make_r_JK_engine(ctx, J, K, P)\nend program\n")
    libdir = os.path.join(ROOT, "tonto_amd")
    r = subprocess.run([FC, "-J", str(tmp_path), src, str(prog), "-o", str(tmp_path / "a.out"),
                        f"-L{libdir}", "-ltonto_gpu", f"-Wl,-rpath,{libdir}",
                        "-Wl,--allow-shlib-undefined"],
                       capture_output=True, text=True, cwd=tmp_path)
    if r.returncode != 0 and "libtonto_gpu" in r.stderr:
        # the linker needs the library under its lib* name
        os.symlink(os.path.join(libdir, "libtonto_gpu.so"), tmp_path / "libtonto_gpu.so")
    assert r.returncode == 0, r.stderr[-2000:]
