"""The drop-in boundary against the reference's own sources (CPU only;
skipped where the reference tree is absent).

* Every function the reference declares in include/{fluid,temperature,
  param,parallel,fileio,statistics,save,logging}.h is declared by the
  header of the same name in include/ with a compatible type: a C file
  includes this build's header, takes the address of each function and then
  repeats the reference's declaration verbatim, which the compiler rejects
  as "conflicting types" if any parameter or return type differs.
* The reference's src/main.c, unchanged, compiles against include/ and
  links against lib/libsns.so + lib/libsns_hip.so (plus the MPI library
  for its MPI_Init / MPI_Finalize): the library supplies every other symbol
  the time loop calls, logging, save and statistics included.  The program
  is linked, not run (running it needs a GPU).
"""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
LIB = os.path.join(ROOT, "simplenavierstokessolver_amd", "lib")
REF = "/root/reference"
HEADERS = ["fluid", "temperature", "param", "parallel", "fileio", "statistics", "save",
           "logging"]

pytestmark = pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "include")),
                                reason="reference tree absent")


def mpi_dirs():
    from simplenavierstokessolver_amd import build
    inc = build.mpi_include()
    lib = os.path.join(os.path.dirname(inc), "lib")
    return inc, lib


def declarations(path):
    """(name, text) of every extern function declaration in a header."""
    src = open(path).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    src = re.sub(r"//[^\n]*", "", src)
    out = []
    for m in re.finditer(r"\bextern\s+[^;{}]*?\(\s*[^;{}]*\)\s*;", src):
        text = " ".join(m.group(0).split())
        name = re.search(r"(\w+)\s*\(", text).group(1)
        out.append((name, text))
    return out


@pytest.mark.parametrize("header", HEADERS)
def test_reference_prototypes_are_declared_identically(header, tmp_path):
    decls = declarations(os.path.join(REF, "include", header + ".h"))
    assert decls, header
    inc, _ = mpi_dirs()
    c = ['#include "%s.h"' % header]
    c.append("void *sns_declared[] = {%s};" % ", ".join("(void *)%s" % n for n, _ in decls))
    c += [t for _, t in decls]
    f = tmp_path / ("check_%s.c" % header)
    f.write_text("\n".join(c) + "\n")
    r = subprocess.run(["gcc", "-std=c99", "-fsyntax-only", "-Werror", "-I", INC, "-I", inc,
                        str(f)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_reference_main_links_against_libsns(tmp_path):
    for so in ("libsns.so", "libsns_hip.so"):
        if not os.path.exists(os.path.join(LIB, so)):
            pytest.skip("libraries not built")
    inc, mpilib = mpi_dirs()
    exe = tmp_path / "a.out.gpu"
    r = subprocess.run(["gcc", "-std=c99", "-O2", "-Wall", "-Wextra", "-I", INC, "-I", inc,
                        os.path.join(REF, "src", "main.c"), "-o", str(exe),
                        "-L" + LIB, "-lsns", "-lsns_hip", "-Wl,-rpath," + LIB,
                        os.path.join(mpilib, "libmpi.so"), "-lm"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    # every symbol the program needs beyond libc / libm / MPI comes from libsns
    nm = subprocess.run(["nm", "-D", "--undefined-only", str(exe)], capture_output=True,
                        text=True, check=True).stdout.split()
    needed = {s for s in nm if re.match(r"^(param|parallel|fluid|temperature|statistics|save|"
                                        r"logging)\w*$", s)}
    exported = subprocess.run(["nm", "-D", "--defined-only", os.path.join(LIB, "libsns.so")],
                              capture_output=True, text=True, check=True).stdout.split()
    assert needed and needed <= set(exported), needed - set(exported)
