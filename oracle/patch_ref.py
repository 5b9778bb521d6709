#!/usr/bin/env python3
"""Build recipe for the reference oracle binary ``oracle/_ref/irisfdtd_ref``.

TEST INFRASTRUCTURE ONLY.  This script is part of the oracle: it is run by
``oracle/Makefile`` in the build container, never by the product path.

It copies the reference Fortran sources from ``<refroot>/code`` into a
temporary directory outside the repository, applies a small portability
patch that LLVM Flang (``amdflang``) needs, compiles them with
``amdflang -O2`` together with ``oracle/ref_dump.f90`` (an opt-in dump of the
on-the-fly DFT accumulators, our own code) and writes only the linked binary
to ``oracle/_ref/``.  The temporary copy is deleted afterwards, so no
reference source lands in the repository or travels with it.

The edits touch no numerics.  They are the Intel-Fortran extensions Flang
rejects:

* INTEGER flags used as LOGICAL conditions (``if(SDL_ON)``) -> ``/=0``
  (code/files.f90:152, code/driver.f90:290,468, code/sdl.f90:368-843,
  code/common.f90:158);
* ``not(logical)`` -> ``.not.(logical)`` (code/driver.f90:247,448,462,519,
  code/dataprocess.f90:341,364);
* the Intel ``EOF()`` intrinsic in the section-search loops
  (code/files.f90:175,258,300) -> ``iostat`` checks;
* ``iostat=`` on section-search reads so that a ``(``-led comment line does
  not abort list-directed input (code/files.f90:51,120,158,324,
  code/sdl.f90:895);
* a bare real literal read into the COMPLEX ``eps_bbox``
  (code/files.f90:164) -> read a real and assign.

Plus one inserted call after each ``scattering_fourier_out`` in
code/driver.f90:460,463 to dump ``emg`` when ``IRIS_REF_DUMP`` is set.
"""
import os
import re
import shutil
import subprocess
import sys
import tempfile

ORDER = ["parameters", "physconst", "files", "common", "sdl", "mesh", "energy",
         "boundary", "boundary_II", "fdtd", "tfsf", "dataprocess", "driver",
         "main", "program"]

EOF_BLOCK = re.compile(
    r"if\(not\(eof\(inFDTD\)\)\)\s*then\s*\n(\s*)read\(inFDTD,\*\)\s*msg\s*\n"
    r"\s*else\s*\n\s*exit\s*\n\s*end if", re.IGNORECASE)

# (file, pattern, replacement, expected match count); patterns are regexes.
EDITS = [
    ("files.f90", r"if\(SDL_plane_on\(1\)\+SDL_plane_on\(2\)\+SDL_plane_on\(3\)\) then",
     "if((SDL_plane_on(1)+SDL_plane_on(2)+SDL_plane_on(3))/=0) then", 1),
    ("files.f90", r"character\(100\) grating,runname",
     "character(100) grating,runname\ninteger :: ios_ref\nreal(8) :: eps_bbox_re", 1),
    ("files.f90", r"read\(inFDTD,\*\) eps_bbox\b",
     "read(inFDTD,*) eps_bbox_re; eps_bbox = eps_bbox_re", 1),
    ("files.f90", r"(do while\(msg/=(?:msg0|msg1|msg5|grating)\)\s*\n\s*)read\(inFDTD,\*\) msg",
     r"\1read(inFDTD,*,iostat=ios_ref) msg", 4),
    ("sdl.f90", r"read\(inputSDL_FILE,\*\) message",
     "read(inputSDL_FILE,*,iostat=ios_ref) message", 1),
    ("sdl.f90", r"if\((SDL_rotate_180|SDL_rotate_90|SDL_copy_flag|SDL_scale)\)",
     r"if(\1/=0)", 16),
    ("sdl.f90", r"if\(\(SDL_copy_flag\)\.or\.\(SDL_rotate_90\)\.or\.\(SDL_scale\)\.or\.\(SDL_rotate_180\)\)",
     "if((SDL_copy_flag/=0).or.(SDL_rotate_90/=0).or.(SDL_scale/=0).or.(SDL_rotate_180/=0))", 2),
    ("driver.f90", r"if\(SDL_ON\)", "if(SDL_ON/=0)", 2),
    ("driver.f90", r"\bnot\(single_box\)", ".not.(single_box)", 4),
    ("dataprocess.f90", r"\bnot\(default\)", ".not.(default)", 2),
    ("common.f90", r"if\(graphic\)then", "if(graphic/=0)then", 1),
    ("driver.f90", r"(call scattering_fourier_out\(emg,spts,it,1\))",
     r"\1; call iris_ref_dump(emg(1)%fdata,size(emg(1)%fdata,1),size(emg(1)%fdata,2),1,1,it,nt);"
     r" call iris_ref_dump(emg(2)%fdata,size(emg(2)%fdata,1),size(emg(2)%fdata,2),1,2,it,nt);"
     r" call iris_ref_dump(emg(3)%fdata,size(emg(3)%fdata,1),size(emg(3)%fdata,2),1,3,it,nt)", 1),
    ("driver.f90", r"(call scattering_fourier_out\(emg2,spts2,it,2\))",
     r"\1; call iris_ref_dump(emg2(1)%fdata,size(emg2(1)%fdata,1),size(emg2(1)%fdata,2),2,1,it,nt);"
     r" call iris_ref_dump(emg2(2)%fdata,size(emg2(2)%fdata,1),size(emg2(2)%fdata,2),2,2,it,nt);"
     r" call iris_ref_dump(emg2(3)%fdata,size(emg2(3)%fdata,1),size(emg2(3)%fdata,2),2,3,it,nt)", 1),
]


def patch_tree(src_dir, dst_dir):
    for name in ORDER:
        with open(os.path.join(src_dir, name + ".f90"), "rb") as f:
            text = f.read().decode("latin-1")
        if name == "files":
            text, n = EOF_BLOCK.subn(r"read(inFDTD,*,iostat=ios_ref) msg\n\1if(ios_ref<0) exit", text)
            if n != 3:
                raise SystemExit(f"patch: EOF blocks in files.f90 matched {n}, expected 3")
        if name == "sdl":
            # sdl's setparam_SDL needs its own iostat variable
            text, n = re.subn(r"(character\(LONGITUD_ARCHIVO\)::message)", r"\1\n  integer :: ios_ref", text)
            if n != 1:
                raise SystemExit("patch: setparam_SDL header not found")
        for fname, pat, rep, count in EDITS:
            if fname != name + ".f90":
                continue
            text, n = re.subn(pat, rep, text)
            if n != count:
                raise SystemExit(f"patch: {fname}: '{pat}' matched {n}, expected {count}")
        with open(os.path.join(dst_dir, name + ".f90"), "wb") as f:
            f.write(text.encode("latin-1"))


def main():
    refroot = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    out = sys.argv[2] if len(sys.argv) > 2 else os.path.join(os.path.dirname(__file__), "_ref", "irisfdtd_ref")
    fc = os.environ.get("FC", "amdflang")
    here = os.path.dirname(os.path.abspath(__file__))
    tmp = tempfile.mkdtemp(prefix="iris_ref_build_")
    try:
        patch_tree(os.path.join(refroot, "code"), tmp)
        shutil.copy(os.path.join(here, "ref_dump.f90"), tmp)
        objs = []
        for name in ["ref_dump"] + ORDER:
            subprocess.run([fc, "-O2", "-c", name + ".f90", "-o", name + ".o"], cwd=tmp, check=True,
                           stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
            objs.append(name + ".o")
        os.makedirs(os.path.dirname(os.path.abspath(out)), exist_ok=True)
        subprocess.run([fc, "-O2", "-o", os.path.abspath(out)] + objs, cwd=tmp, check=True)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    print("built", out)


if __name__ == "__main__":
    main()
