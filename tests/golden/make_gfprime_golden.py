"""Writes the pole-file fixtures of tests/golden:

gfprime_flang.restart   unformatted, written by gfprime_writer.f90 (defined
                        WRITE(UNFORMATTED) as in VCA_AUX_FUNX.f90:566-582,
                        one parent WRITE per element as in :633-647),
                        compiled here with the image's Fortran compiler;
gfprime_gfortran_style.restart
                        formatted, the same contents laid out as gfortran's
                        list-directed output of write_formatted
                        (VCA_AUX_FUNX.f90:484-506): integers in 12
                        columns, complex(8) as (  x.xxxxxxxxxxxxxxxxx     ,
                        ...) with 17 significant digits, the \\n character
                        item after each state, one line per element.

The expected contents follow from golden_gf_contents() (the formula in the
header of gfprime_writer.f90).  Run: python tests/golden/make_gfprime_golden.py
"""
from __future__ import annotations

import os
import shutil
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
FC = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"


def golden_gf_contents():
    """{(ilat, jlat, 1, 1, 1): [state][channel] -> (weights, poles)} with
    complex values, Nlat=2, Nspin=1, Norb=1."""
    out = {}
    k = 0
    for il in (1, 2):
        for jl in (1, 2):
            k += 1
            states = []
            for s in (1, 2):
                chans = []
                for c in range(1, (2 if il == jl else 4) + 1):
                    n = (k + s + c) % 4
                    poles = [complex(k + s / 4 + c / 16 + j / 256, 0.0) for j in range(1, n + 1)]
                    w = [complex(j / 8, -(s + c) / 32) for j in range(1, n + 1)]
                    chans.append((w, poles))
                states.append(chans)
            out[(il, jl, 1, 1, 1)] = states
    return out


def _g(x: float) -> str:
    """gfortran's list-directed real(8) field (G25.17 with a 3-digit
    exponent), e.g. '  1.5000000000000000     ' or '  1.0000000000000000E-003'."""
    if x == 0.0:
        return "  0.0000000000000000     "
    ax = abs(x)
    sign = "-" if x < 0 else " "
    if 0.1 <= ax < 1e16:
        digits = 17 if ax < 1 else 16 - len(str(int(ax))) + 1
        return f" {sign}{ax:.{digits}f}     "
    m, e = f"{ax:.16E}".split("E")
    return f" {sign}{m}E{int(e):+04d}"


def write_gfortran_style(path):
    lines = []
    for key in sorted(golden_gf_contents()):
        states = golden_gf_contents()[key]
        items = [f"{len(states):12d}"]
        for chans in states:
            items.append(f"{len(chans):12d}")
            for (w, poles) in chans:
                items.append(f"{len(poles):12d}")
                items += [f"({_g(z.real)},{_g(z.imag)})" for z in poles]
                items += [f"({_g(z.real)},{_g(z.imag)})" for z in w]
            items.append(" \\n")
        lines.append("".join(" " + it if not it.startswith(" ") else it for it in items))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def write_flang(path):
    with tempfile.TemporaryDirectory() as td:
        exe = os.path.join(td, "gfprime_writer")
        subprocess.check_call([FC, "-module-dir", td, os.path.join(HERE, "gfprime_writer.f90"),
                               "-o", exe])
        subprocess.check_call([exe], cwd=td)
        shutil.copyfile(os.path.join(td, "gfprime_flang.restart"), path)


if __name__ == "__main__":
    write_flang(os.path.join(HERE, "gfprime_flang.restart"))
    write_gfortran_style(os.path.join(HERE, "gfprime_gfortran_style.restart"))
    print("wrote", os.listdir(HERE))
