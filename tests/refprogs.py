"""The reference test suite's example programs, written in this package's
DSL (same rules, names and plans as the reference's
tests/test_integration_*.py / test_hir_mir_tc_e2e.py), paired with the MIR
goldens those tests check (tests/golden/mir/)."""

from srdatalog_hip.dsl import SPLIT, Filter, PlanEntry, Program, Relation, Var


def tc() -> Program:
  X, Y, Z = Var("x"), Var("y"), Var("z")
  arc, edge, path = Relation("ArcInput", 2), Relation("Edge", 2), Relation("Path", 2)
  return Program(rules=[
    (edge(X, Y) <= arc(X, Y)).named("EdgeLoad"),
    (path(X, Y) <= edge(X, Y)).named("TCBase"),
    (path(X, Z) <= path(X, Y) & edge(Y, Z)).named("TCRec"),
  ])


def tc_2level_bg() -> Program:
  X, Y, Z = Var("x"), Var("y"), Var("z")
  arc, edge, path = Relation("ArcInput", 2), Relation("Edge", 2), Relation("Path", 2)
  return Program(rules=[
    (edge(X, Y) <= arc(X, Y)).named("EdgeLoad"),
    (path(X, Y) <= edge(X, Y)).named("TCBase"),
    (path(X, Z) <= path(X, Y) & edge(Y, Z)).named("TCRec")
    .with_plan(delta=0, block_group=True, var_order=["y", "x", "z"]),
  ])


def sg() -> Program:
  P, Q, X, Y = Var("p"), Var("q"), Var("x"), Var("y")
  arc, sgr = Relation("Arc", 2), Relation("Sg", 2)
  return Program(rules=[
    (sgr(X, Y) <= arc(P, X) & arc(P, Y) & Filter(vars=("x", "y"), code="return x != y;"))
    .named("SGBase").with_plan(var_order=["p", "x", "y"]),
    (sgr(X, Y) <= arc(P, X) & sgr(P, Q) & arc(Q, Y))
    .named("SGRec").with_plan(var_order=["p", "q", "x", "y"]),
  ])


def andersen() -> Program:
  X, Y, Z, W = Var("x"), Var("y"), Var("z"), Var("w")
  ao, asgn = Relation("AddressOf", 2), Relation("Assign", 2)
  load_r, store_r, pt = Relation("Load", 2), Relation("Store", 2), Relation("PointsTo", 2)
  return Program(rules=[
    (pt(Y, X) <= ao(Y, X)).named("Base").with_plan(var_order=["y", "x"]),
    (pt(Y, X) <= pt(Z, X) & asgn(Y, Z)).named("Assign")
    .with_plan(delta=0, clause_order=[0, 1], var_order=["z", "x", "y"]),
    (pt(Y, W) <= pt(X, Z) & load_r(Y, X) & pt(Z, W)).named("Load").with_plans([
      PlanEntry(delta=0, clause_order=(0, 1, 2), var_order=("x", "z", "y", "w")),
      PlanEntry(delta=2, clause_order=(2, 0, 1), var_order=("z", "x", "y", "w")),
    ]),
    (pt(Z, W) <= pt(Y, Z) & store_r(Y, X) & pt(X, W)).named("Store").with_plans([
      PlanEntry(delta=0, clause_order=(0, 1, 2), var_order=("y", "x", "z", "w")),
      PlanEntry(delta=2, clause_order=(2, 1, 0), var_order=("x", "y", "w", "z")),
    ]).with_inject_cpp("debug"),
  ])


def lsqb_triangle_count() -> Program:
  A, B, C, X, Y = Var("a"), Var("b"), Var("c"), Var("x"), Var("y")
  k_in, knows, tri = Relation("KnowsInput", 2), Relation("Knows", 2), Relation("Triangle", 3)
  return Program(rules=[
    (knows(X, Y) <= k_in(X, Y)).named("KnowsLoad"),
    (tri(A, B, C) <= knows(A, B) & knows(B, C) & knows(A, C)).named("TriangleJoin")
    .with_plan(var_order=["a", "b", "c"]).with_count(),
  ])


def triangle() -> Program:
  X, Y, Z, H, F = Var("x"), Var("y"), Var("z"), Var("h"), Var("f")
  rr, sr, tr, zr = Relation("RRel", 2), Relation("SRel", 3), Relation("TRel", 3), Relation("ZRel", 3)
  return Program(rules=[
    (zr(X, Y, Z) <= rr(X, Y) & sr(Y, Z, H) & tr(Z, X, F)).named("Triangle")
    .with_plan(var_order=["x", "y", "z"]),
  ])


def lsqb_q9_neg2hop() -> Program:
  """examples/lsqb_q9_neg2hop.py: 2-hop paths closed by no direct edge."""
  P, P1, P2, P3, T, X, Y = (Var(n) for n in ("p", "p1", "p2", "p3", "t", "x", "y"))
  k_in, hi_in = Relation("KnowsInput", 2), Relation("HasInterestInput", 2)
  knows, hi, path = Relation("Knows", 2), Relation("HasInterest", 2), Relation("Path", 4)
  return Program(rules=[
    (knows(X, Y) <= k_in(X, Y)).named("KnowsLoad"),
    (hi(P, T) <= hi_in(P, T)).named("InterestLoad"),
    (path(P1, P2, P3, T) <= knows(P1, P2) & knows(P2, P3) & hi(P3, T) & ~knows(P1, P3)
     & Filter(("p1", "p3"), "return p1 != p3;"))
    .named("NegPath").with_plan(var_order=["p2", "p3", "p1", "t"]).with_count(),
  ])


def lsqb_q7_optional() -> Program:
  """examples/lsqb_q7_optional.py: the four OPTIONAL-MATCH cases, two of
  them with negated marker relations."""
  C, CM, L, M_, T, U = (Var(n) for n in ("c", "comment", "liker", "m", "t", "_"))
  has_tag, has_creator = Relation("HasTag", 2), Relation("HasCreator", 2)
  reply_of, likes = Relation("ReplyOf", 2), Relation("Likes", 2)
  has_reply, has_liker = Relation("HasReply", 1), Relation("HasLiker", 1)
  c1, c2 = Relation("Case1", 5), Relation("Case2", 4)
  c3, c4 = Relation("Case3", 4), Relation("Case4", 3)
  return Program(rules=[
    (has_reply(M_) <= reply_of(U, M_)).named("MarkHasReply"),
    (has_liker(M_) <= likes(Var("_"), M_)).named("MarkHasLiker"),
    (c1(M_, T, C, CM, L) <= has_tag(M_, T) & has_creator(M_, C) & reply_of(CM, M_) & likes(L, M_))
    .named("Case1Both").with_plan(var_order=["m", "comment", "liker", "t", "c"]).with_count(),
    (c2(M_, T, C, CM) <= has_tag(M_, T) & has_creator(M_, C) & reply_of(CM, M_) & ~has_liker(M_))
    .named("Case2ReplyOnly").with_plan(var_order=["m", "comment", "t", "c"]).with_count(),
    (c3(M_, T, C, L) <= has_tag(M_, T) & has_creator(M_, C) & likes(L, M_) & ~has_reply(M_))
    .named("Case3LikeOnly").with_plan(var_order=["m", "liker", "t", "c"]).with_count(),
    (c4(M_, T, C) <= has_tag(M_, T) & has_creator(M_, C) & ~has_reply(M_) & ~has_liker(M_))
    .named("Case4Neither").with_plan(var_order=["m", "t", "c"]).with_count(),
  ])


def lsqb_q3_triangle() -> Program:
  """examples/lsqb_q3_triangle.py: triangles whose three persons live in
  cities of one country."""
  V = {n: Var(n) for n in ("c", "c1", "c2", "c3", "co", "p", "p1", "p2", "p3", "x", "y")}
  k_in, loc_in, part_in = (Relation("KnowsInput", 2), Relation("IsLocatedInInput", 2),
                           Relation("IsPartOfInput", 2))
  knows, loc, part = Relation("Knows", 2), Relation("IsLocatedIn", 2), Relation("IsPartOf", 2)
  tri = Relation("Triangle", 3)
  x, y, p, c, co = V["x"], V["y"], V["p"], V["c"], V["co"]
  p1, p2, p3, c1, c2, c3 = V["p1"], V["p2"], V["p3"], V["c1"], V["c2"], V["c3"]
  return Program(rules=[
    (knows(x, y) <= k_in(x, y)).named("KnowsLoad"),
    (loc(p, c) <= loc_in(p, c)).named("LocLoad"),
    (part(c, co) <= part_in(c, co)).named("PartOfLoad"),
    (tri(p1, p2, p3) <= knows(p1, p2) & knows(p2, p3) & knows(p3, p1) & loc(p1, c1) & loc(p2, c2)
     & loc(p3, c3) & part(c1, co) & part(c2, co) & part(c3, co))
    .named("LabeledTriangle").with_plan(var_order=["p1", "p2", "p3", "c1", "c2", "c3", "co"]),
  ])


def _q6(two_sources: bool, rev: bool, head: str, rule: str, order, count=False, bg=False):
  P, P1, P2, P3, T, X, Y = (Var(n) for n in ("p", "p1", "p2", "p3", "t", "x", "y"))
  k_in, hi_in = Relation("KnowsInput", 2), Relation("HasInterestInput", 2)
  knows, hi, path = Relation("Knows", 2), Relation("HasInterest", 2), Relation(head, 4)
  rules = [(knows(X, Y) <= k_in(X, Y)).named("KnowsLoad")]
  second = knows
  if rev:
    rules.append((knows(Y, X) <= k_in(X, Y)).named("KnowsLoadRev"))
  if two_sources:
    k2_in, second = Relation("KnowsInput2", 2), Relation("Knows2", 2)
    rules.append((second(X, Y) <= k2_in(X, Y)).named("Knows2Load"))
  rules.append((hi(P, T) <= hi_in(P, T)).named("InterestLoad"))
  r = (path(P1, P2, P3, T) <= knows(P1, P2) & second(P2, P3) & hi(P3, T)
       & Filter(("p1", "p3"), "return p1 != p3;")).named(rule).with_plan(var_order=order,
                                                                         block_group=bg)
  if count:
    r = r.with_count()
  rules.append(r)
  return Program(rules=rules)


def lsqb_q6_2hop() -> Program:
  return _q6(False, False, "Path", "TwoHopPath", ["p1", "p2", "p3", "t"])


def lsqb_q6_count() -> Program:
  return _q6(False, True, "Path", "TwoHopPath", ["p2", "p3", "p1", "t"], count=True)


def lsqb_q6_nosj() -> Program:
  return _q6(True, False, "Path", "TwoHopPath", ["p1", "p2", "p3", "t"])


def lsqb_q6_bg_test() -> Program:
  return _q6(True, False, "PathBG", "TwoHopBG", ["p2", "p3", "p1", "t"], bg=True)


def split_test() -> Program:
  """Q(x, z) :- A(x, y), !B(y), split, C(x, z) (the reference's
  tests/fixtures/split_test.nim): negation pushed below a split."""
  X, Y, Z = Var("x"), Var("y"), Var("z")
  a, b, c, q = Relation("A", 2), Relation("B", 1), Relation("C", 2), Relation("Q", 2)
  return Program(rules=[(q(X, Z) <= a(X, Y) & ~b(Y) & SPLIT & c(X, Z)).named("SplitTest")])


def split_rec() -> Program:
  """R(x, y) :- Seed(x, y);  R(x, z) :- R(x, y), !Filter(y), split, Edge(y, z)
  (the reference's tests/fixtures/split_rec.nim)."""
  X, Y, Z = Var("x"), Var("y"), Var("z")
  seed, filt = Relation("Seed", 2), Relation("Filter", 1)
  edge, r = Relation("Edge", 2), Relation("R", 2)
  return Program(rules=[
    (r(X, Y) <= seed(X, Y)).named("RBase"),
    (r(X, Z) <= r(X, Y) & ~filt(Y) & SPLIT & edge(Y, Z)).named("RRec"),
  ])


GOLDEN = {
  "tc": tc, "tc_2level_bg": tc_2level_bg, "sg": sg, "andersen": andersen,
  "lsqb_triangle_count": lsqb_triangle_count, "triangle": triangle,
  "lsqb_q9_neg2hop": lsqb_q9_neg2hop, "lsqb_q7_optional": lsqb_q7_optional,
  "lsqb_q3_triangle": lsqb_q3_triangle, "lsqb_q6_2hop": lsqb_q6_2hop,
  "lsqb_q6_count": lsqb_q6_count, "lsqb_q6_nosj": lsqb_q6_nosj,
  "lsqb_q6_bg_test": lsqb_q6_bg_test, "split_test": split_test, "split_rec": split_rec,
}
