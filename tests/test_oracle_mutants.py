"""The oracle's pins reject plausible mistakes.

Each case copies oracle/swe_oracle.c, plants one mistake a restatement of
wind.hpp / world.hpp could plausibly contain, builds the copy with the
oracle's own flags into a temporary directory, and runs the pin suites
against it (SWE_ORACLE_LIB).  The suites must fail.  Everything runs on the
CPU; the planted mistakes never leave the temporary directory.
"""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "swe_oracle.c")
PINS = ["tests/test_oracle_fly_pins.py", "tests/test_oracle_march.py", "tests/test_oracle_geometry.py",
        "tests/test_oracle_primitives.py", "tests/test_oracle_colour.py"]

MUTANTS = {
    # wind.hpp:66 -- gravity with the wrong sign
    "gravity_sign": ("if (airborne) sp.y -= gravity * sed;", "if (airborne) sp.y += gravity * sed;"),
    # wind.hpp:83 -- turbulence a tenth of its size
    "turbulence_scale": ("float k = 0.1f * hfac * collision;", "float k = 0.01f * hfac * collision;"),
    # wind.hpp:87 -- the drag factor computed in float
    "float_drag": ("const float drag = (float)(1.0 - 0.3 * (double)sed);", "const float drag = 1.0f - 0.3f * sed;"),
    # wind.hpp:75 -- the outer cross product transposed
    "redirect_transposed": ("v3 rspeed = cross3(n, cross3(smul(1.0f - collision, sp), n));",
                            "v3 rspeed = cross3(cross3(smul(1.0f - collision, sp), n), n);"),
    # wind.hpp:79 -- mix's operands swapped
    "mix_swapped": ("mix3(ps, rspeed, shadow * hfac)", "mix3(rspeed, ps, shadow * hfac)"),
    # wind.hpp:50 -- hfac from the wrong sign of the height difference
    "hfac_sign": ("float hfac = or_expf(-(w->py - h) / p->boundarylayer);",
                  "float hfac = or_expf((w->py - h) / p->boundarylayer);"),
    # wind.hpp:104 -- lift with 1 + collision
    "lift_sign": ("const float lift = (1.0f - collision) * len3(sp);",
                  "const float lift = (1.0f + collision) * len3(sp);"),
    # world.hpp:139 -- the sequential cascade gives only half to the neighbour
    "seq_cascade_half": ("      c->height -= t; nc->height += t;", "      c->height -= t; nc->height += 0.5f * t;"),
    # colour-ordered cascade: the classes walked y class major instead of x class major
    "colour_class_order": ("  for (int32_t cx = 0; cx < ncx; cx++)\n    for (int32_t cy = 0; cy < ncy; cy++) {",
                           "  for (int32_t cy = 0; cy < ncy; cy++)\n    for (int32_t cx = 0; cx < ncx; cx++) {"),
    # colour-ordered cascade: later classes read the neighbours' heights as they were
    "colour_stale_neighbours": ("          for (int32_t k = 0; k < cnt; k++) refresh(p, nbc[k], zm, sm, S, SS);",
                                "          (void)nbc;"),
    # wind.hpp:111 -- the sequential pickup added to the cell
    "seq_pickup_sign": ("    cell->height -= p->suspension * diff;\n    if (p->layers == 2) {",
                        "    cell->height += p->suspension * diff;\n    if (p->layers == 2) {"),
}


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_pins_reject_planted_mistake(name, tmp_path):
    old, new = MUTANTS[name]
    text = open(SRC).read()
    assert text.count(old) == 1, "the mutated line is no longer in the oracle"
    src = tmp_path / "swe_oracle.c"
    src.write_text(text.replace(old, new))
    lib = tmp_path / "liboracle.so"
    subprocess.run(["gcc", "-O2", "-std=gnu11", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-fopenmp",
                    "-shared", "-o", str(lib), str(src), "-lm"], check=True)
    env = dict(os.environ, SWE_ORACLE_LIB=str(lib), OMP_NUM_THREADS="4")
    res = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", *PINS],
                         cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    assert res.returncode == 1, "the pins accepted the mistake %r:\n%s" % (name, res.stdout[-2000:])
