"""Mutation check of the oracle's pins (TEST INFRASTRUCTURE ONLY).

Builds the oracle source with one planted mistake at a time into a temporary
directory, runs the CPU pins (tests/test_oracle.py and
tests/test_oracle_closed_forms.py) against that build through REYN_ORACLE_LIB,
and reports whether they fail, as they must. Nothing is written into the
repository. Usage:

    python oracle/mutations.py            # every mutant
    python oracle/mutations.py visc_10nu  # named ones

Each mutant names the reference lines whose term it breaks.
"""

from __future__ import annotations

import os
import subprocess
import sys
import tempfile
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent
SRC = HERE / "reyn_oracle.cpp"
TESTS = [str(ROOT / "tests" / "test_oracle.py"), str(ROOT / "tests" / "test_oracle_closed_forms.py")]

# name: (reference term, text in the source, replacement, occurrence (0-based) of the text)
MUTANTS = {
    "density_no_boundary": ("boundary density, solver.cuh:34",
                            "rho[i] = acc + accb;", "rho[i] = acc;", 0),
    "visc_10nu": ("viscosity factor 10 nu, solver.cuh:62",
                  "const T c = T(10) * T(p.nu);", "const T c = T(2) * T(p.nu);", 0),
    "visc_eps": ("viscosity 0.01 h^2, solver.cuh:63",
                 "const T eps = T(0.01) * T(p.h) * T(p.h);", "const T eps = T(0.1) * T(p.h) * T(p.h);", 0),
    "visc_mi_over_rhoi": ("viscosity m_j / rho_j, solver.cuh:62",
                          "c * s.m[j] / rho[j] * dot3(vij, d)", "c * s.m[i] / rho[i] * dot3(vij, d)", 0),
    "sesph_visc_10nu": ("SESPH viscosity factor 10 nu, SESPH.cu:50",
                        "const T c = T(10) * T(p.nu);", "const T c = T(2) * T(p.nu);", 1),
    "sesph_eos_unnormalised": ("SESPH EOS k (rho / rho0 - 1), SESPH.cu:9",
                               "kstiff * (r / rho0 - T(1))", "kstiff * (r - rho0)", 0),
    "sesph_boundary_factor": ("SESPH boundary pressure -(p_i / rho_i^2), SESPH.cu:57-61",
                              "ay[i] = a[1] - pri * gb[1];", "ay[i] = a[1] - T(0.5) * pri * gb[1];", 0),
    "sesph_boundary_dropped": ("SESPH boundary pressure term, SESPH.cu:57-61",
                               "ax[i] = a[0] - pri * gb[0];\n        ay[i] = a[1] - pri * gb[1];\n"
                               "        az[i] = a[2] - pri * gb[2];",
                               "ax[i] = a[0];\n        ay[i] = a[1];\n        az[i] = a[2];", 0),
    "pcisph_boundary_half": ("PCISPH boundary factor 1/2, PCISPH.cu:49",
                             "const T cb = T(0.5) * (pri + pr0);", "const T cb = (pri + pr0);", 0),
    "pcisph_boundary_rho0": ("PCISPH boundary p_i / rho0^2 term, PCISPH.cu:49",
                             "const T cb = T(0.5) * (pri + pr0);", "const T cb = T(0.5) * (pri + pri);", 0),
    "pcisph_eos_normalised": ("PCISPH EOS k (rho - rho0), PCISPH.cu:18",
                              "const T pi = std::fmax(T(0), kk * (ri - rho0));",
                              "const T pi = std::fmax(T(0), kk * (ri / rho0 - T(1)));", 0),
    "predict_fluid_velocity": ("predicted density dt grad W . v_ij, solver.cuh:96-98",
                               "acc += s.m[j] * (kW(k, d2) + dt * dot3(dw, vij));", "acc += s.m[j] * kW(k, d2);", 0),
    "predict_boundary_velocity": ("predicted density boundary dt grad W . v_i, solver.cuh:100-104",
                                  "acc += b->m[j] * (kW(k, d2) + dt * dot3(dw, vi));", "acc += b->m[j] * kW(k, d2);",
                                  0),
    "iisph_aii_no_boundary": ("IISPH a_ii boundary gradient sum, IISPH.cu:48-61",
                              "const T Gs[3] = { gf[0] + G[0], gf[1] + G[1], gf[2] + G[2] };",
                              "const T Gs[3] = { gf[0], gf[1], gf[2] };", 0),
    "iisph_aii_no_cross_term": ("IISPH a_ii square of the combined sum, IISPH.cu:53-61",
                                "const T a = -dt * (dot3(Gs, Gs) + s.m[i] * sq) / (rho0 * rho0);",
                                "const T a = -dt * (dot3(gf, gf) + dot3(G, G) + s.m[i] * sq) / (rho0 * rho0);", 0),
    "iisph_source_no_boundary": ("IISPH source boundary divergence, IISPH.cu:63-65",
                                 "(div + dot3(vi, G))", "(div)", 0),
    "iisph_source_sign": ("IISPH source (rho0 - rho) / dt, IISPH.cu:65",
                          "const T sv = (rho0 - rho[i]) / dt", "const T sv = (rho[i] - rho0) / dt", 0),
    "iisph_pacc_no_boundary": ("IISPH pressure acceleration boundary term, IISPH.cu:84-95",
                               "ay[i] = -r0sqinv * (a[1] + pi * G[1]);", "ay[i] = -r0sqinv * (a[1]);", 0),
    "iisph_Ap_no_boundary": ("IISPH Ap boundary term, IISPH.cu:142-147",
                             "lap += dot3(ai, G);", "", 0),
    "iisph_Ap_no_dt": ("IISPH Ap factor dt, IISPH.cu:136",
                       "const T Ap = dt * lap;", "const T Ap = lap;", 0),
}


def mutate(text: str, old: str, new: str, occurrence: int) -> str:
    pos = -1
    for _ in range(occurrence + 1):
        pos = text.find(old, pos + 1)
        if pos < 0:
            raise ValueError(f"text not found (occurrence {occurrence}): {old!r}")
    return text[:pos] + new + text[pos + len(old):]


def build(src: Path, out: Path) -> None:
    """The oracle's own build recipe (oracle/Makefile) on another source."""
    subprocess.run(["make", "-s", "-B", "-C", str(HERE), f"SRC={src}", f"OUT={out}"], check=True,
                   capture_output=True, text=True)


def run_pins(lib: Path) -> tuple[int, str]:
    env = dict(os.environ, REYN_ORACLE_LIB=str(lib), OMP_NUM_THREADS="2")
    proc = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", *TESTS],
                          cwd=str(ROOT), env=env, capture_output=True, text=True)
    failed = [ln.split(" ")[1] for ln in proc.stdout.splitlines() if ln.startswith("FAILED ")]
    return proc.returncode, failed[0] if failed else ""


def main(names: list[str]) -> int:
    text = SRC.read_text()
    names = names or list(MUTANTS)
    bad = 0
    with tempfile.TemporaryDirectory() as tmp:
        tmp = Path(tmp)
        # the unmutated source passes
        build(SRC, tmp / "base.so")
        rc, _ = run_pins(tmp / "base.so")
        print(f"{'unmutated':28s} pins {'pass' if rc == 0 else 'FAIL'}")
        bad += rc != 0
        for name in names:
            term, old, new, occ = MUTANTS[name]
            src = tmp / f"{name}.cpp"
            src.write_text(mutate(text, old, new, occ))
            build(src, tmp / f"{name}.so")
            rc, first = run_pins(tmp / f"{name}.so")
            caught = rc != 0
            bad += not caught
            print(f"{name:28s} {'caught' if caught else 'NOT CAUGHT':10s} {term}; first failure: {first}")
    return 1 if bad else 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
