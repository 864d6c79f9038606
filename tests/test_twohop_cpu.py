"""LSQB Q6 (the reference's examples/lsqb_q6_2hop.py), host side: the
brute-force model against a hand-worked case and against the closed form
for |Path| from sparse matrix products (self-loops included, which the
p1 != p3 filter and the diagonal term must both see); the same for the
model of what the example's plan order compiles to (p1 a successor of p2,
as the reference's MIR for it reads; see programs.twohop_program); and the
program's compilation to a generated join with the filter on (p1, p3)."""

import numpy as np
import pytest

import oracle.brute as brute


def inputs(seed, n, m, tags, mi):
  """Random Knows pairs (self-loops allowed) and HasInterest pairs."""
  rng = np.random.default_rng(seed)
  k = np.unique(rng.integers(0, n, size=(m, 2)).astype(np.uint32), axis=0)
  h = np.unique(np.stack([rng.integers(0, n, mi), rng.integers(0, tags, mi)], 1)
                .astype(np.uint32), axis=0)
  return k, h


def test_hand_case():
  # 0->1->2 (tag 7), 0->1->0 (filtered: p1 == p3), 1->1->2 via the loop
  knows = [(0, 1), (1, 2), (1, 0), (1, 1)]
  interest = [(2, 7), (0, 5), (1, 9)]
  rels, deltas = brute.twohop(knows, interest)
  assert rels["Path"] == {(0, 1, 2, 7), (1, 1, 2, 7), (1, 1, 0, 5), (0, 1, 1, 9)}
  assert deltas["Path"] == [4]
  assert brute.twohop_count_closed_form(knows, interest, 3) == 4
  # as planned: p1, p3 distinct successors of p2.  succ(1) = {0, 1, 2}:
  # (p1, p3) in {(0,1), (0,2), (1,0), (1,2), (2,0), (2,1)}; succ(0) = {1}
  planned = brute.twohop_as_planned(knows, interest)
  assert planned == {(0, 1, 1, 9), (0, 1, 2, 7), (1, 1, 0, 5), (1, 1, 2, 7), (2, 1, 0, 5),
                     (2, 1, 1, 9)}
  assert brute.twohop_as_planned_count_closed_form(knows, interest, 3) == 6


@pytest.mark.parametrize("seed,n,m,tags,mi", [(0, 30, 120, 5, 40), (1, 200, 1500, 12, 300),
                                              (2, 8, 40, 3, 20)])
def test_model_matches_closed_form(seed, n, m, tags, mi):
  k, h = inputs(seed, n, m, tags, mi)
  assert (k[:, 0] == k[:, 1]).any() or n > 8    # the small dense case holds self-loops
  rels, _ = brute.twohop(k, h)
  assert len(rels["Path"]) == brute.twohop_count_closed_form(k, h, n)
  assert all(p1 != p3 for p1, _, p3, _ in rels["Path"])
  planned = brute.twohop_as_planned(k, h)
  assert len(planned) == brute.twohop_as_planned_count_closed_form(k, h, n)
  assert planned != rels["Path"]


def test_closed_form_catches_a_dropped_filter():
  """The closed form separates the model from one without p1 != p3."""
  k, h = inputs(3, 10, 60, 4, 30)
  rels, _ = brute.twohop(k, h)
  kn = {(int(a), int(b)) for a, b in k}
  tags = {}
  for p, t in h:
    tags.setdefault(int(p), set()).add(int(t))
  unfiltered = {(a, b, c, t) for (a, b) in kn for (b2, c) in kn if b2 == b for t in tags.get(c, ())}
  assert len(unfiltered) != brute.twohop_count_closed_form(k, h, 10)
  assert len(rels["Path"]) == brute.twohop_count_closed_form(k, h, 10)


@pytest.mark.parametrize("name", ["twohop", "twohop_count"])
def test_program_compiles(name):
  from srdatalog_hip import codegen, compiler, mir as M, programs
  prog = programs.PROGRAMS[name]()
  mir = compiler.compile_to_mir(prog)
  src = codegen.generate_program_source(name, compiler.relation_decls(prog), mir)
  assert "TwoHopPath" in src and "p1 != p3" in src
  # the listing keeps the example's order, so p2 joins on Knows (0 1)'s first
  # column; the count's p2-first order reads Knows(p1, p2) through (1 0)
  text = M.print_program(mir)
  assert ("(column-source :index (Knows 1 0)" in text) == (name == "twohop_count")
