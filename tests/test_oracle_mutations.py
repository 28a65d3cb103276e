"""Plants plausible misreadings of the reference in a copy of the oracle's C
source and checks that the oracle's CPU pinning tests catch each one.

Each mutation is a textual edit of oracle/rich_oracle.c; the copy is built
with the oracle's own compiler flags into a temporary directory and the
pinning tests run against it in a child process (RICH_ORACLE_LIB).  A
mutation that left every pinning test green would mean that the oracle's
rule it changes is pinned by nothing.  The unmodified source, built the
same way, must pass.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "rich_oracle.c")
PIN_TESTS = ["tests/test_oracle_pins.py", "tests/test_oracle_scheme.py",
             "tests/test_oracle_hllc.py", "tests/test_oracle_pcm.py",
             "tests/test_oracle_moving.py", "tests/test_oracle_limits.py"]

# name -> [(text in rich_oracle.c, replacement)], each text found exactly once
MUTATIONS = {
    # shocked_slope_limit (LinearGaussImproved.cpp:343-366)
    "shocked guard with abs": [("if (fabs(dphi) > 1e-9 * cell) {",
                                "if (fabs(dphi) > fabs(1e-9 * cell)) {")],
    "shocked gate 0.05": [("if (fabs(dphi) > 0.1 * smax(", "if (fabs(dphi) > 0.05 * smax(")],
    "theta 1": [("smax(theta * (nval - cell) / dphi, 0.0)", "smax(1.0 * (nval - cell) / dphi, 0.0)")],
    "shocked limiter skips mirror faces": [
        ("  for (int i = 0; i < n; ++i) {\n    const cc_t* t = &ncell[i];\n",
         "  for (int i = 0; i < n; ++i) {\n"
         "    if (orig_index(m, m->face_first[faces[i]]) == orig_index(m, m->face_second[faces[i]]))"
         " continue;\n    const cc_t* t = &ncell[i];\n")],
    # slope_limit (LinearGaussImproved.cpp:174-300)
    "slope_limit gate 0.1": [("if (fabs(dphi) > 0.05 * smax(", "if (fabs(dphi) > 0.1 * smax(")],
    "velocity guard without abs": [
        ("cell->vx, maxd.vx, mind.vx, 1, &psi[2]", "cell->vx, maxd.vx, mind.vx, 0, &psi[2]"),
        ("cell->vy, maxd.vy, mind.vy, 1, &psi[3]", "cell->vy, maxd.vy, mind.vy, 0, &psi[3]")],
    "no mirror skip": [
        ("    if (orig_index(m, m->face_first[f]) == orig_index(m, m->face_second[f])) continue;\n",
         "")],
    # calc_naive_slope's neighbour limit (LinearGaussImproved.cpp:73-78)
    "neighbour limit at 20": [("if (n > 20) return RO_TOO_MANY_NEIGHBORS;",
                               "if (n >= 20) return RO_TOO_MANY_NEIGHBORS;")],
    # is_shock (LinearGaussImproved.cpp:150-157)
    "no divergence test": [("const int cond1 = (res->x.vx + res->y.vy) * width < (-prm->delta_v) * cs;",
                            "const int cond1 = 0 * (width + cs) != 0;")],
    "no pressure test": [("const int cond2 = pressure_ratio(&cell, n, ncell) < prm->delta_p;",
                          "const int cond2 = 0 * pressure_ratio(&cell, n, ncell) != 0;")],
    # RigidWallGenerator (RigidWallGenerator.cpp:5-37)
    "no rigid reflection": [("      t.vx = t.vx - nx * s2;\n      t.vy = t.vy - ny * s2;\n",
                             "      (void)s2;\n")],
    "rigid ghost from GetOriginalIndex": [("cc_t t = get_cell(c, side == 2 ? a : b);",
                                           "cc_t t = get_cell(c, orig_index(m, g));")],
    "rigid ghost from its last face": [("    if (filled[g - N]) continue;\n", "")],
    # ModularFluxCalculator over PCM (PCM.cpp:5-37)
    "PCM sides swapped in the flux": [
        ("/* ModularFluxCalculator over PCM (modular_flux_calculator.cpp:39-71) */\n"
         "    for (int32_t f = 0; f < F; ++f) {\n"
         "      put_cell(left, f, get_cell(cng, m->face_first[f]));\n"
         "      put_cell(right, f, get_cell(cng, m->face_second[f]));",
         "/* ModularFluxCalculator over PCM (modular_flux_calculator.cpp:39-71) */\n"
         "    for (int32_t f = 0; f < F; ++f) {\n"
         "      put_cell(left, f, get_cell(cng, m->face_second[f]));\n"
         "      put_cell(right, f, get_cell(cng, m->face_first[f]));")],
    # RoundCells (round_cells.cpp:14-76) and CalcFaceVelocity (VoronoiMesh.cpp:146-151)
    "calc_dw threshold eta R": [("if (!(dist < 0.9 * eta * R))", "if (!(dist < eta * R))")],
    "calc_dw without max(c, |v|)": [("smax(dp2c(gamma, d[i], p[i]), vabs(vx[i], vy[i]))",
                                     "dp2c(gamma, d[i], p[i])")],
    "overshoot never turns back": [("      if (R * 0.1 > dists[k]) factor *= -0.1;\n", "")],
    "face velocity without the bisector term": [
        ("fwx[f] = 0.5 * (wlx + wrx) + s * dx / a2;", "fwx[f] = 0.5 * (wlx + wrx) + 0 * s * dx / a2;"),
        ("fwy[f] = 0.5 * (wly + wry) + s * dy / a2;", "fwy[f] = 0.5 * (wly + wry) + 0 * s * dy / a2;")],
    "mirror faces move": [("    if (bc == 0 && oa == ob) {\n      fwx[f] = 0.0;",
                           "    if (bc == 0 && oa == ob && 0) {\n      fwx[f] = 0.0;")],
    # PeriodicGhostGenerator::GetGhostGradient (PeriodicGhostGenerator.cpp:19-24)
    "zero periodic ghost gradient": [("  if (prm->bc == 0) return z;\n", "  return z;\n")],
}


def _build(tmp_path, edits):
    src = open(SRC).read()
    for old, new in edits:
        assert src.count(old) == 1, old
        src = src.replace(old, new)
    c = tmp_path / "rich_oracle.c"
    c.write_text(src)
    lib = tmp_path / "liboracle.so"
    subprocess.run(["gcc", "-O2", "-std=c99", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
                    "-fopenmp", "-I", os.path.join(ROOT, "oracle"), "-shared", "-o", str(lib),
                    str(c), "-lm"], check=True)
    return lib


def _run_pins(lib):
    env = dict(os.environ, RICH_ORACLE_LIB=str(lib))
    return subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider",
                           *PIN_TESTS], cwd=ROOT, env=env, capture_output=True, text=True,
                          timeout=300)


def test_unmodified_oracle_passes(tmp_path):
    r = _run_pins(_build(tmp_path, []))
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]


@pytest.mark.parametrize("name", sorted(MUTATIONS))
def test_mutation_is_caught(tmp_path, name):
    r = _run_pins(_build(tmp_path, MUTATIONS[name]))
    assert r.returncode != 0, f"planted mistake '{name}' passes every pinning test"
    assert "failed" in r.stdout, r.stdout[-3000:] + r.stderr[-2000:]
