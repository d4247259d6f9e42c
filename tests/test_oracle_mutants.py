"""The oracle's pins must catch plausible mistakes.  Each case compiles a copy
of oracle/rf_oracle.c with one planted mistake (a textual substitution that
must match exactly once) into a temporary library and runs the pin that is
meant to catch it: the pin must fail on the mutant and pass on the oracle.
The mutants cover the nearest-hit tie rule and the bounding-surface logic of
both arithmetic forms (RayFactorCL/RayFactorGPU.cl:445-472, :578-624)."""
import importlib.util
import shutil
import subprocess
from pathlib import Path

import pytest

import oracle_pins
from conftest import ROOT

SRC = ROOT / "oracle" / "rf_oracle.c"

FIND32 = "int32_t rfo_find_f32("
FIND64 = "int32_t rfo_find_f64("
BV32 = "int32_t rfo_find_bv_f32("
BV64 = "int32_t rfo_find_bv_f64("

# name: (function whose body is mutated, old text, new text, pin, pin arguments)
MUTANTS = {
    "find_f32_tie_le": (FIND32, "if (t < tbest)", "if (t <= tbest)", "tie", ("compat",)),
    "find_f64_tie_le": (FIND64, "if (t < tbest)", "if (t <= tbest)", "tie", ("fp64",)),
    "bv_f32_stale_index_dropped": (BV32, "if (bvIndex != -1 && objectIndex == -1)", "if (bvIndex != -1)",
                                   "stale", ("compat",)),
    "bv_f64_stale_index_dropped": (BV64, "if (bvIndex != -1 && objectIndex == -1)", "if (bvIndex != -1)",
                                   "stale", ("fp64",)),
    "bv_f32_intermediate_always_taken": (BV32, "if (objectIndex != -1) { tbest = t; bvIndex = w; }",
                                         "{ tbest = t; bvIndex = w; }", "intermediate", ("compat",)),
    "bv_f64_intermediate_always_taken": (BV64, "if (objectIndex != -1) { tbest = t; bvIndex = w; }",
                                         "{ tbest = t; bvIndex = w; }", "intermediate", ("fp64",)),
    "bv_f64_child_range_short": (BV64, "bvs[bvIndex].objEndIndex, ray)", "bvs[bvIndex].objEndIndex - 1, ray)",
                                 "closed", ("input-C14-BV", "fp64")),
    "bv_f32_child_range_short": (BV32, "bvs[bvIndex].objEndIndex, ray)", "bvs[bvIndex].objEndIndex - 1, ray)",
                                 "closed", ("input-C52-BV", "compat")),
    "bv_f64_surface_far_wins": (BV64, "if (t < tbest) {", "if (t > tbest || tbest == INFINITY) {",
                                "stale", ("fp64",)),
}


def mutate(src: str, func: str, old: str, new: str) -> str:
    start = src.index(func)
    end = src.index("\n}\n", start)
    body = src[start:end]
    assert body.count(old) == 1, f"{old!r} must occur once in {func}"
    return src[:start] + body.replace(old, new) + src[end:]


def load_oracle(lib: Path):
    """A separate instance of oracle/oracle.py bound to another library."""
    spec = importlib.util.spec_from_file_location(f"oracle_{lib.stem}", ROOT / "oracle" / "oracle.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    mod.LIB_PATH = lib
    mod.lib()
    return mod


def run_pin(O, kind, args, golden_scene):
    if kind == "tie":
        oracle_pins.pin_find_tie(O, *args)
    elif kind == "stale":
        oracle_pins.pin_bv_stale_index(O, *args)
    elif kind == "intermediate":
        oracle_pins.pin_bv_intermediate_miss(O, *args)
    else:
        name, mode = args
        oracle_pins.pin_bv_closed_form(O, name, golden_scene(name), mode)


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_pin_catches_mutant(tmp_path, oracle, golden_scene, name):
    func, old, new, kind, args = MUTANTS[name]
    src = mutate(SRC.read_text(), func, old, new)
    (tmp_path / "rf_oracle.c").write_text(src)
    shutil.copy(ROOT / "oracle" / "rf_oracle.h", tmp_path / "rf_oracle.h")
    lib = tmp_path / f"librf_oracle_{name}.so"
    subprocess.run(["gcc", "-O2", "-march=x86-64-v3", "-ffp-contract=off", "-fno-math-errno", "-fPIC", "-std=c11",
                    "-shared", "-o", str(lib), str(tmp_path / "rf_oracle.c"), "-lm", "-lpthread"], check=True)
    run_pin(oracle, kind, args, golden_scene)            # the oracle passes the pin
    mutant = load_oracle(lib)
    with pytest.raises(AssertionError):
        run_pin(mutant, kind, args, golden_scene)        # the mutant fails it
