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


def _fortran_program(mol, P):
    """A Fortran main program that fills tgpu_system for `mol`, builds J and K of P
    through the module (the make_r_JK_engine replacement), writes them to jk.txt, and
    reports a failing call's message (negative cutoff) the way the shim would."""
    a = mol.shell_arrays()

    def arr(name, typ, vals, fmt):
        body = ", ".join(fmt(v) for v in vals)
        return f"  {typ}, target :: {name}({len(vals)}) = [{body}]\n"
    d = lambda v: f"{float(v):.17e}_c_double"  # noqa: E731
    i = lambda v: f"{int(v)}_c_int"  # noqa: E731
    n = mol.n_bf
    src = "program run_build\n  use tonto_gpu_mod\n  implicit none\n"
    src += arr("xyz", "real(c_double)", a["atom_xyz"].ravel(), d)
    src += arr("zz", "real(c_double)", a["atom_Z"], d)
    src += arr("sl", "integer(c_int)", a["l"], i)
    src += arr("sa", "integer(c_int)", a["atom"], i)
    src += arr("sn", "integer(c_int)", a["nprim"], i)
    src += arr("ex", "real(c_double)", a["exps"], d)
    src += arr("co", "real(c_double)", a["coefs"], d)
    src += arr("pv", "real(c_double)", P.ravel(order="F"), d)
    src += (f"  real(c_double) :: P({n},{n}), J({n},{n}), K({n},{n}), maxI({mol.n_shell * (mol.n_shell + 1) // 2})\n"
            "  type(tgpu_system) :: sys\n  type(c_ptr) :: ctx\n  integer(c_int) :: rc\n  integer :: u\n"
            f"  P = reshape(pv, [{n}, {n}])\n"
            f"  sys%n_atom = {mol.n_atom}\n  sys%n_shell = {mol.n_shell}\n"
            "  sys%atom_xyz = c_loc(xyz)\n  sys%atom_Z = c_loc(zz)\n  sys%shell_l = c_loc(sl)\n"
            "  sys%shell_atom = c_loc(sa)\n  sys%shell_nprim = c_loc(sn)\n  sys%exps = c_loc(ex)\n"
            "  sys%coefs = c_loc(co)\n"
            "  rc = tgpu_create(sys, ctx)\n  if (rc /= TGPU_OK) call tgpu_die(ctx, rc, 'create')\n"
            "  rc = tgpu_set_cutoffs(ctx, -1.0_c_double, 1.0e-9_c_double, 1.0e-9_c_double, 1.0e-6_c_double)\n"
            "  write(*, '(a, i0, a, a)') 'negative cutoff: ', rc, ' ', tgpu_error_message(ctx)\n"
            "  rc = tgpu_set_cutoffs(ctx, 1.0e-9_c_double, 1.0e-9_c_double, 1.0e-9_c_double, 1.0e-6_c_double)\n"
            "  rc = tgpu_schwarz(ctx, maxI)\n  if (rc /= TGPU_OK) call tgpu_die(ctx, rc, 'schwarz')\n"
            "  call gpu_make_r_JK_engine(ctx, J, K, P)\n"
            "  open(newunit=u, file='jk.txt')\n  write(u, '(es25.17)') J\n  write(u, '(es25.17)') K\n"
            "  close(u)\n  rc = tgpu_destroy(ctx)\nend program\n")
    return src


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(FC), reason="no Fortran compiler in this image")
def test_fortran_program_builds_jk(tmp_path, oracle_mod):
    """The Fortran side end to end on the GPU: iso_c_binding module, tgpu_create from
    Fortran arrays, tgpu_schwarz, the make_r_JK_engine replacement body; J and K against
    the oracle (1e-10), and a failing call's message read through tgpu_error_message."""
    import numpy as np
    from tonto_amd.geometry import tonto_water
    mol = tonto_water("6-31G(d)")
    rng = np.random.default_rng(3)
    A = rng.standard_normal((mol.n_bf, mol.n_bf))
    P = A @ A.T / mol.n_bf
    prog = tmp_path / "run_build.f90"
    prog.write_text(_fortran_program(mol, P))
    src = os.path.join(ROOT, "tonto_amd", "fortran", "tonto_gpu_mod.f90")
    libdir = os.path.join(ROOT, "tonto_amd")
    r = subprocess.run([FC, "-J", str(tmp_path), src, str(prog), "-o", str(tmp_path / "run_build"),
                        f"-L{libdir}", "-ltonto_gpu", f"-Wl,-rpath,{libdir}"],
                       capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 0, r.stderr[-3000:]
    out = subprocess.run([str(tmp_path / "run_build")], capture_output=True, text=True, cwd=tmp_path,
                         timeout=120)
    print(out.stdout[-2000:], out.stderr[-2000:])
    assert out.returncode == 0
    assert "negative cutoff: 1 cutoffs must be non-negative" in out.stdout
    vals = np.loadtxt(tmp_path / "jk.txt")
    n = mol.n_bf
    J = vals[:n * n].reshape(n, n, order="F")
    K = vals[n * n:].reshape(n, n, order="F")
    o = oracle_mod.OracleSystem(mol)
    Jo, Ko = o.make_r_JK(P)
    assert np.max(np.abs(J - Jo)) < 1e-10 and np.max(np.abs(K - Ko)) < 1e-10
