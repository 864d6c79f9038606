"""The DSL -> MIR compiler reproduces the reference's MIR goldens byte for
byte (tests/golden/mir/*.mir.sexpr are the reference's own fixtures), and
the MIR printer/parser round-trips them."""

import glob
import os

import pytest

import refprogs
from srdatalog_hip import codegen, compiler, mir, programs

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mir")


@pytest.mark.parametrize("name", sorted(refprogs.GOLDEN))
def test_compiler_matches_golden(name):
  got = mir.print_program(compiler.compile_to_mir(refprogs.GOLDEN[name]()))
  want = open(os.path.join(GOLD, f"{name}.mir.sexpr")).read().rstrip("\n")
  assert got == want


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLD, "*.mir.sexpr"))))
def test_print_parse_roundtrip(path):
  text = open(path).read().rstrip("\n")
  assert mir.print_program(mir.parse_program(text)) == text


def test_benchmark_programs_match_reference_plans():
  # The benchmark programs carry input_file / print_size pragmas but plan
  # exactly like the reference's test programs of the same rules.
  same = {"tc": "tc", "sg": "sg", "triangle_count": "lsqb_triangle_count"}
  for prog, gold in same.items():
    got = mir.print_program(compiler.compile_to_mir(programs.PROGRAMS[prog]()))
    want = open(os.path.join(GOLD, f"{gold}.mir.sexpr")).read().rstrip("\n")
    assert got == want, prog


def test_andersen_plan_matches_golden_modulo_hooks():
  got = compiler.compile_to_mir(programs.PROGRAMS["andersen"]())
  want = mir.parse_program(open(os.path.join(GOLD, "andersen.mir.sexpr")).read())
  strip = lambda p: [st for st in p.steps]  # noqa: E731
  # the benchmark program has no debug hook; drop the golden's hook nodes
  for st in strip(want):
    if isinstance(st.body, mir.FixpointPlan):
      st.body.ops = [o for o in st.body.ops if not isinstance(o, mir.InjectCppHook)]
  assert mir.print_program(got) == mir.print_program(want)


@pytest.mark.parametrize("name", sorted(programs.PROGRAMS))
def test_codegen_emits_every_program(name):
  prog = programs.PROGRAMS[name]()
  src = codegen.generate_program_source(name, compiler.relation_decls(prog),
                                        compiler.compile_to_mir(prog))
  assert "srdatalog_init" in src and "k_pipes" in src


def test_stratify_orders_dependencies():
  from srdatalog_hip.dsl import Program, Relation, Var
  x, y = Var("x"), Var("y")
  a, b, c = Relation("A", 2), Relation("B", 2), Relation("C", 2)
  p = Program(rules=[(c(x, y) <= b(x, y)).named("R2"), (b(x, y) <= a(x, y)).named("R1")])
  st = compiler.stratify(list(p.rules))
  order = [sorted(s.members) for s in st]
  assert order.index(["B"]) < order.index(["C"])
