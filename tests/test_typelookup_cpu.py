"""Constants in rules (heads, body atoms, negations) and the type-lookup
test program: the brute-force model pinned on a hand-worked case, the
compiler's binding of constants, and the generated code's literals."""

import numpy as np

import oracle.brute as brute
from srdatalog_hip import codegen, compiler, programs
from srdatalog_hip.dsl import Const, Relation, Var


def test_brute_typelookup_hand_case():
  # classes 10 <- 11 <- 12 (12's superclass is 11, 11's is 10); methods:
  # 20 in 10 named 5; 21 in 11 named 5, abstract; 22 in 12 named 7
  # (main), static; 23 in 12 named 6, static
  sup = [(11, 10), (12, 11)]
  cls = [(10,), (11,), (12,)]
  decl = [(20, 10), (21, 11), (22, 12), (23, 12)]
  name = [(20, 5), (21, 5), (22, 7), (23, 6)]
  mod = [(1, 21), (2, 22), (2, 23)]
  out = brute.typelookup(sup, cls, decl, name, mod, [(12,)])
  assert out["Implemented"] == {(10, 5, 20), (12, 7, 22), (12, 6, 23)}   # 21 is abstract
  # 11 and 12 inherit 10's name 5 (none of their own); 12's own names stay
  assert out["Lookup"] == {(10, 5, 20), (11, 5, 20), (12, 5, 20), (12, 7, 22), (12, 6, 23)}
  assert out["MainMethod"] == {(22,)}
  assert out["Initialized"] == {(12,), (11,), (10,)}
  assert out["SubtypeOf"] == {(10, 10), (11, 11), (12, 12), (11, 10), (12, 11), (12, 10),
                              (10, 0), (11, 0), (12, 0)}


def test_brute_typelookup_filter_excludes_the_initialiser():
  # method 3 (the class initialiser's id) is named main and static: the
  # filter keeps it out
  out = brute.typelookup([], [(1,)], [(3, 1)], [(3, 7)], [(2, 3)], [(1,)])
  assert out["MainMethod"] == set() and out["Initialized"] == set()


def test_constants_bind_as_filters_and_literals():
  prog = programs.PROGRAMS["typelookup"]()
  mir = compiler.compile_to_mir(prog)
  text = str(mir)
  # body constants: a fresh variable and an equality filter each
  assert 'return K1_7 == 7u;' in text and 'return K2_2 == 2u;' in text
  src = codegen.generate_program_source("typelookup", compiler.relation_decls(prog), mir)
  # negated constant: a literal leading the prefix; wildcard: only the bound prefix
  assert "!srdj::has_prefix(V.v[2], 1u, v_m)" in src
  # head constant: written as a literal, and its rule writes rows (no mark mode)
  i = src.index("_SubRoot {")
  body = src[i:src.index("}  // namespace", i)]
  assert "= 0u;" in body and "run_k<6" not in body


def test_const_values_are_u32():
  r = Relation("R", 2)
  x = Var("x")
  assert r(x, 5).args == ("x", "#5") and r(x, Const(7)).args == ("x", "#7")
  for bad in (-1, 1 << 32, 1.5, True):
    try:
      r(x, bad)
    except (TypeError, ValueError):
      continue
    raise AssertionError(f"{bad!r} accepted")


def test_brute_cspa_hand_case():
  # assign 1 <- 2, deref: 1 points to 3 and 2 points to 4
  out, deltas = brute.cspa([(1, 2)], [(1, 3), (2, 4)])
  vf = {(1, 2), (1, 1), (2, 2)}
  assert out["ValueFlow"] == vf
  # value aliases: pairs with a common value-flow source (1 -> {1, 2}, 2 -> {2})
  assert out["ValueAlias"] == {(1, 1), (1, 2), (2, 1), (2, 2)}
  # memory aliases: reflexive on 1 and 2, and the dereferences of aliased 1, 2
  assert out["MemoryAlias"] == {(1, 1), (2, 2), (3, 3), (3, 4), (4, 3), (4, 4)}
  assert deltas["ValueFlow"][-1] == deltas["ValueAlias"][-1] == deltas["MemoryAlias"][-1] == 0
  # the program compiles to one recursive stratum holding all three relations
  from srdatalog_hip import compiler
  strata = compiler.stratify(compiler.rename_wildcards(list(programs.PROGRAMS["cspa"]().rules)))
  rec = [s for s in strata if s.recursive]
  assert any({"ValueFlow", "ValueAlias", "MemoryAlias"} <= set(s.members) for s in rec)


def test_brute_galen_hand_case():
  # P: 1->2, 2->3; Q: (3, 7, 4); R: (7, 8, 9) with Q(4, 8, 5) gives
  # Q(x, 9, 5) for every x reaching 4 through Q(x, 7, 4); S: 7 -> 6
  p = [(1, 2), (2, 3)]
  q = [(3, 7, 4), (4, 8, 5)]
  out, deltas = brute.galen(p, q, [(7, 8, 9)], [], [], [(7, 6)])
  assert out["OutP"] == {(1, 2), (2, 3), (1, 3)}
  # PropQ carries Q(3, ., .) back along P to 2 and 1; PropQS renames 7 -> 6;
  # Join3c: Q(x, 7, 4) and Q(4, 8, 5) with R(7, 8, 9) give Q(x, 9, 5)
  qx = {(3, 7, 4), (4, 8, 5), (3, 6, 4), (3, 9, 5)}
  want = qx | {(a, r, c) for a in (1, 2) for (_, r, c) in qx if _ == 3}
  assert out["OutQ"] == want
  assert deltas["OutP"][-1] == deltas["OutQ"][-1] == 0
