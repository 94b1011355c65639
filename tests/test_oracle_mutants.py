"""The oracle pins catch plausible mistakes (CPU).

Each case below plants one mistake in a copy of oracle/oracle_body.h (a swapped corner, a
wrong stride or clamp, a rounded index, a dropped factor, a uniform-spacing assumption),
builds that copy with the oracle's own compiler flags into a temporary directory, and runs
the pin that is meant to catch it against the copy: the pin must fail there.  The planted
text is matched exactly once, so a change to the oracle that removes it shows up here.
"""

import concurrent.futures as cf
import os
import re
import subprocess

import numpy as np  # noqa: F401  (the pins use it)
import pytest

import test_oracle_pins as P

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE = os.path.join(os.path.dirname(HERE), "oracle")

# (name, text in oracle_body.h, replacement, pins that must fail)
MUTANTS = [
    ("map: corner 2 takes corner 3's index",
     "idx[2] = idx[1] + g->mp[0];", "idx[2] = idx[0] + g->mp[0];", ["pin_map_lookup"]),
    ("map: the max-face clamp one cell too low",
     "if (t[d] >= (R)(g->mp[d] - 1)) t[d] = (R)(g->mp[d] - 2);",
     "if (t[d] >= (R)(g->mp[d] - 1)) t[d] = (R)(g->mp[d] - 3);", ["pin_map_lookup"]),
    ("map: no clamp below the origin",
     "if (t[d] < (R)0) t[d] = (R)0;", "", ["pin_map_lookup"]),
    ("map: slice stride from the wrong axis",
     "idx[4] = idx[0] + g->mp[0] * g->mp[1];", "idx[4] = idx[0] + g->mp[0] * g->mp[2];",
     ["pin_map_lookup"]),
    ("map: parametric coordinate over two cells",
     "R hi = g->morg[d] + g->msp[d] * (R)(cell[d] + 1);",
     "R hi = g->morg[d] + g->msp[d] * (R)(cell[d] + 2);", ["pin_map_lookup"]),
    ("map: estimate returns opacity, not transmittance",
     "R tr = (R)1 - o;", "R tr = o;", ["pin_map_lookup"]),
    ("cell sampler: cell id with point dims",
     "int64_t cid = (cell[2] * g->c[1] + cell[1]) * g->c[0] + cell[0];",
     "int64_t cid = (cell[2] * g->n[1] + cell[1]) * g->n[0] + cell[0];",
     ["pin_sample_cell_entries", "pin_sample_cell_walk"]),
    ("cell sampler: rounded TF index",
     "int64_t ci = (int64_t)(ns * (R)cms);", "int64_t ci = (int64_t)(ns * (R)cms + (R)0.5);",
     ["pin_sample_cell_entries", "pin_sample_cell_walk"]),
    ("cell sampler: TF index clamped one entry short of the top",
     "      if (ci > cms) ci = cms;\n      for (int c = 0; c < 4; ++c) scol[c] = g->tf[4 * ci + c];",
     "      if (ci > cms) ci = cms - 1;\n      for (int c = 0; c < 4; ++c) scol[c] = g->tf[4 * ci + c];",
     ["pin_sample_cell_entries"]),
    ("cell sampler: alpha not scaled by the remaining transparency",
     "R a = scol[3] * ((R)1 - color[3]);", "R a = scol[3];",
     ["pin_sample_cell_constant", "pin_sample_cell_walk"]),
    ("cell sampler: no final clamp",
     "for (int c = 0; c < 4; ++c) color[c] = FN(vmin)(color[c], (R)1);", "",
     ["pin_sample_cell_constant"]),
    ("locator: reciprocal width from uniform spacing",
     "invsp[d] = (R)1 / (hi - lo);",
     "invsp[d] = (R)(g->n[d] - 1) / (g->maxp[d] - g->minp[d]);",
     ["pin_locate_cell_uneven", "pin_sample_cell_walk"]),
    ("point sampler: bottom-left corner from the next coordinate",
     "      bl[0] = g->x[0][cell[0]];\n      bl[1] = g->x[1][cell[1]];\n      bl[2] = g->x[2][cell[2]];\n      s0 = field[i0];",
     "      bl[0] = g->x[0][cell[0] + 1];\n      bl[1] = g->x[1][cell[1] + 1];\n      bl[2] = g->x[2][cell[2] + 1];\n      s0 = field[i0];",
     ["pin_locate_cell_uneven"]),
    ("point sampler: x lerp weighted by ty",
     "R l01 = s0 + tx * s10;", "R l01 = s0 + ty * s10;", ["pin_locate_cell_uneven"]),
]


def _cflags():
    """The oracle's compiler and flags, read from its Makefile."""
    mk = open(os.path.join(ORACLE, "Makefile")).read()
    flags = re.search(r"^CFLAGS\s*:=\s*((?:.*\\\n)*.*)$", mk, re.M).group(1)
    return flags.replace("\\\n", " ").split()


def _build(tmp, k, old, new):
    body = open(os.path.join(ORACLE, "oracle_body.h")).read()
    assert body.count(old) == 1, f"planted text not found exactly once: {old!r}"
    d = os.path.join(tmp, f"m{k}")
    os.makedirs(d)
    with open(os.path.join(d, "oracle_body.h"), "w") as f:
        f.write(body.replace(old, new))
    for name in ("oracle.c", "oracle.h"):
        with open(os.path.join(ORACLE, name)) as src, open(os.path.join(d, name), "w") as dst:
            dst.write(src.read())
    out = os.path.join(d, "liboracle.so")
    subprocess.run(["gcc", *_cflags(), "-shared", "-o", out, os.path.join(d, "oracle.c"), "-lm"],
                   check=True, capture_output=True)
    return out


@pytest.fixture(scope="module")
def mutant_libs(tmp_path_factory):
    tmp = str(tmp_path_factory.mktemp("oracle_mutants"))
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 2)) as ex:
        futs = [ex.submit(_build, tmp, k, old, new) for k, (_, old, new, _) in enumerate(MUTANTS)]
        return [f.result() for f in futs]


def test_oracle_passes_every_pin(oracle_mod):
    for pin in sorted({p for m in MUTANTS for p in m[3]}):
        getattr(P, pin)(oracle_mod)


@pytest.mark.parametrize("k", range(len(MUTANTS)), ids=[m[0] for m in MUTANTS])
def test_pin_catches_planted_mistake(oracle_mod, mutant_libs, k):
    O = oracle_mod
    name, _, _, pins = MUTANTS[k]
    saved = O._lib
    O._lib = O.load(mutant_libs[k])
    try:
        for pin in pins:
            with pytest.raises(AssertionError):
                getattr(P, pin)(O)
    finally:
        O._lib = saved
