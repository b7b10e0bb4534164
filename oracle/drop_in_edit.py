"""Writes a reference Solver source with the drop-in edits applied to stdout,
to be piped into the compiler (no copy of a reference file is written).

The edits are the ones a maintainer makes to call the MI355X core from the
reference's own Fortran host (the handle module is
nemoh_amd/fortran/nemoh_core_host.f90, the interfaces
nemoh_amd/fortran/nemoh_core_bind.f90):

* Solver/NEMOH.f90: USE NEMOH_CORE_HOST; CORE_START after INITIALIZE
  (line 84); CORE_STOP before DEALLOCATE_DATA (line 174).
* Solver/Core/SOLVE_BEM_INFD_DIRECT.f90: the influence matrix, GAUSSZ
  inverse, sources and potentials (lines 81-163) become one
  nemoh_core_solve call; the set-up before them (GM, CINT_INFD, N=NQ) stays.
* Solver/Core/SOLVE_BEM_FD_DIRECT.f90: the same for lines 102-183, after
  CINT_FD and LISC.

Every edit checks the text of the lines it touches and fails on a
different source.  Usage: python3 drop_in_edit.py <reference root> <file>
"""
import sys
from pathlib import Path

SOLVE_INFD = ["    CALL CORE_CHECK(nemoh_core_solve(CORE, T, 0.D0, 0, AMBDA, AR, NQ, QQ, CQ, &",
              "         NVEL, ZIGB, ZIGS, ZPB, ZPS))"]
SOLVE_FD = ["    CALL CORE_CHECK(nemoh_core_solve(CORE, T, AMH, NEXP, AMBDA, AR, NQ, QQ, CQ, &",
            "         NVEL, ZIGB, ZIGS, ZPB, ZPS))"]

# (kind, first line, last line, text expected in the first line, text
# expected in the last line, new lines); 1-based inclusive line numbers
EDITS = {
    "Solver/NEMOH.f90": [
        ("insert_after", 40, 40, "USE INITIALIZATION", None, ["    USE NEMOH_CORE_HOST"]),
        ("insert_after", 84, 84, "CALL INITIALIZE(", None, ["    CALL CORE_START()"]),
        ("insert_before", 174, 174, "CALL DEALLOCATE_DATA", None, ["    CALL CORE_STOP()"]),
    ],
    "Solver/Core/SOLVE_BEM_INFD_DIRECT.f90": [
        ("insert_after", 29, 29, "USE M_SOLVER", None, ["  USE NEMOH_CORE_BIND"]),
        ("replace", 81, 163, "Construction of the influence matrix", "END DO", SOLVE_INFD),
    ],
    "Solver/Core/SOLVE_BEM_FD_DIRECT.f90": [
        ("insert_after", 29, 29, "USE M_SOLVER", None, ["  USE NEMOH_CORE_BIND"]),
        ("replace", 102, 183, "Construction of the influence matrix", "END DO", SOLVE_FD),
    ],
}


def edited(ref_root: str, rel: str) -> bytes:
    lines = Path(ref_root, rel).read_bytes().decode("latin-1").split("\n")
    # bottom-up, so earlier line numbers stay valid
    for kind, a, b, first, last, new in sorted(EDITS[rel], key=lambda e: -e[1]):
        if first not in lines[a - 1] or (last is not None and last not in lines[b - 1]):
            raise SystemExit(f"{rel}: lines {a}-{b} are not the expected ones")
        if kind == "insert_after":
            lines[a:a] = new
        elif kind == "insert_before":
            lines[a - 1:a - 1] = new
        else:
            lines[a - 1:b] = new
    return "\n".join(lines).encode("latin-1")


if __name__ == "__main__":
    sys.stdout.buffer.write(edited(sys.argv[1], sys.argv[2]))
