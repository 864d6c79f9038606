"""The benchmark programs, written in the DSL as the reference's examples
write them (examples/tc.py, examples/sg.py, examples/andersen.py,
examples/lsqb_triangle_count.py), plus single-source reachability.

``build_all()`` compiles every program into ``_lib/programs/lib<name>.so``;
``load(name)`` returns its ctypes binding (runtime.ProgramLib).
"""

from __future__ import annotations

import os
import shutil
from concurrent.futures import ThreadPoolExecutor

from . import build as B
from . import native
from .dsl import SPLIT, Const, Filter, PlanEntry, Program, Relation, Var
from .runtime import ProgramLib

PROG_DIR = os.path.join(native.LIB_DIR, "programs")


def tc_program() -> Program:
  """Transitive closure (examples/tc.py)."""
  x, y, z = Var("x"), Var("y"), Var("z")
  arc = Relation("ArcInput", 2, input_file="Arc.csv")
  edge = Relation("Edge", 2)
  path = Relation("Path", 2, print_size=True)
  return Program(rules=[
    (edge(x, y) <= arc(x, y)).named("EdgeLoad"),
    (path(x, y) <= edge(x, y)).named("TCBase"),
    (path(x, z) <= path(x, y) & edge(y, z)).named("TCRec"),
  ])


def sg_program() -> Program:
  """Same generation (examples/sg.py)."""
  p, q, x, y = Var("p"), Var("q"), Var("x"), Var("y")
  arc = Relation("Arc", 2, input_file="Arc.csv")
  sg = Relation("Sg", 2, print_size=True)
  return Program(rules=[
    (sg(x, y) <= arc(p, x) & arc(p, y) & Filter(("x", "y"), "return x != y;"))
    .named("SGBase").with_plan(var_order=["p", "x", "y"]),
    (sg(x, y) <= arc(p, x) & sg(p, q) & arc(q, y))
    .named("SGRec").with_plan(var_order=["p", "q", "x", "y"]),
  ])


def andersen_program() -> Program:
  """Andersen points-to (examples/andersen.py; the benchmark's four rules
  with variables renamed)."""
  w, x, y, z = Var("w"), Var("x"), Var("y"), Var("z")
  addr = Relation("AddressOf", 2, input_file="addressOf.csv")
  assign = Relation("Assign", 2, input_file="assign.csv")
  load = Relation("Load", 2, input_file="load.csv")
  store = Relation("Store", 2, input_file="store.csv")
  pt = Relation("PointsTo", 2, print_size=True)
  return Program(rules=[
    (pt(y, x) <= addr(y, x)).named("Base").with_plan(var_order=["y", "x"]),
    (pt(y, x) <= pt(z, x) & assign(y, z)).named("Assign")
    .with_plan(var_order=["z", "x", "y"]),
    (pt(y, w) <= pt(x, z) & load(y, x) & pt(z, w)).named("Load")
    .with_plan(delta=0, var_order=["x", "z", "y", "w"])
    .with_plan(delta=2, var_order=["z", "x", "y", "w"]),
    (pt(z, w) <= pt(y, z) & store(y, x) & pt(x, w)).named("Store")
    .with_plan(delta=0, var_order=["y", "x", "z", "w"])
    .with_plan(delta=2, var_order=["x", "y", "w", "z"]),
  ])


def triangle_program(count_only: bool) -> Program:
  """Triangle listing / count (examples/lsqb_triangle_count.py).  With the
  input oriented low->high, every result has a < b < c."""
  a, b, c, x, y = Var("a"), Var("b"), Var("c"), Var("x"), Var("y")
  k_in = Relation("KnowsInput", 2, input_file="Person_knows_Person.csv")
  knows = Relation("Knows", 2)
  tri = Relation("Triangle", 3, print_size=not count_only)
  rule = (tri(a, b, c) <= knows(a, b) & knows(b, c) & knows(a, c)).named("TriangleJoin") \
      .with_plan(var_order=["a", "b", "c"])
  if count_only:
    rule = rule.with_count()
  return Program(rules=[(knows(x, y) <= k_in(x, y)).named("KnowsLoad"), rule])


def reach_program() -> Program:
  """Single-source reachability: Reach(y) <= Src(y); Reach(z) <= Reach(y), Edge(y,z)."""
  y, z = Var("y"), Var("z")
  src = Relation("Src", 1, input_file="Src.csv")
  edge = Relation("Edge", 2, input_file="Edge.csv")
  reach = Relation("Reach", 1, print_size=True)
  return Program(rules=[
    (reach(y) <= src(y)).named("ReachBase"),
    (reach(z) <= reach(y) & edge(y, z)).named("ReachRec"),
  ])


def neg2hop_program(count_only: bool) -> Program:
  """LSQB Q9, two-hop paths not closed by a direct edge
  (examples/lsqb_q9_neg2hop.py; the reference rule is count-only, the
  listing variant materialises Path for tuple-level parity)."""
  P, P1, P2, P3, T, X, Y = (Var(n) for n in ("p", "p1", "p2", "p3", "t", "x", "y"))
  k_in = Relation("KnowsInput", 2, input_file="Person_knows_Person.csv")
  hi_in = Relation("HasInterestInput", 2, input_file="Person_hasInterest_Tag.csv")
  knows, hi = Relation("Knows", 2), Relation("HasInterest", 2)
  path = Relation("Path", 4, print_size=not count_only)
  rule = (path(P1, P2, P3, T) <= knows(P1, P2) & knows(P2, P3) & hi(P3, T) & ~knows(P1, P3)
          & Filter(("p1", "p3"), "return p1 != p3;")) \
      .named("NegPath").with_plan(var_order=["p2", "p3", "p1", "t"])
  if count_only:
    rule = rule.with_count()
  return Program(rules=[
    (knows(X, Y) <= k_in(X, Y)).named("KnowsLoad"),
    (hi(P, T) <= hi_in(P, T)).named("InterestLoad"),
    rule,
  ])


def twohop_program(count_only: bool, p2_first: bool | None = None) -> Program:
  """LSQB Q6, two-hop paths ending at a person with an interest
  (examples/lsqb_q6_2hop.py: TwoHopPath over Knows, Knows, HasInterest with
  p1 != p3, plan order p1, p2, p3, t).  p1 occurs in one atom only, so it is
  no join variable, and that order keeps it ahead of p2 in Knows(p1, p2)'s
  index (0 1): the compiled plan, the reference's own
  (tests/fixtures/integration/lsqb_q6_2hop.mir.sexpr), joins p2 on that
  column and reads p1 as a successor of p2.  The count variant takes the
  reference's count test's order p2, p3, p1, t
  (tests/test_integration_lsqb_q6_count.py), whose Knows(p1, p2) index is
  (1 0): it counts the rule as written."""
  P, P1, P2, P3, T, X, Y = (Var(n) for n in ("p", "p1", "p2", "p3", "t", "x", "y"))
  k_in = Relation("KnowsInput", 2, input_file="Person_knows_Person.csv")
  hi_in = Relation("HasInterestInput", 2, input_file="Person_hasInterest_Tag.csv")
  knows, hi = Relation("Knows", 2), Relation("HasInterest", 2)
  path = Relation("Path", 4, print_size=True)
  rule = (path(P1, P2, P3, T) <= knows(P1, P2) & knows(P2, P3) & hi(P3, T)
          & Filter(("p1", "p3"), "return p1 != p3;")) \
      .named("TwoHopPath").with_plan(
          var_order=["p2", "p3", "p1", "t"] if (count_only if p2_first is None else p2_first)
          else ["p1", "p2", "p3", "t"])
  if count_only:
    rule = rule.with_count()
  return Program(rules=[
    (knows(X, Y) <= k_in(X, Y)).named("KnowsLoad"),
    (hi(P, T) <= hi_in(P, T)).named("InterestLoad"),
    rule,
  ])


def q3_program() -> Program:
  """LSQB Q3, directed triangles whose persons live in one country
  (examples/lsqb_q3_triangle.py): a seven-variable, six-level join."""
  V = {n: Var(n) for n in ("c", "c1", "c2", "c3", "co", "p", "p1", "p2", "p3", "x", "y")}
  k_in = Relation("KnowsInput", 2, input_file="Person_knows_Person.csv")
  loc_in = Relation("IsLocatedInInput", 2, input_file="Person_isLocatedIn_City.csv")
  part_in = Relation("IsPartOfInput", 2, input_file="City_isPartOf_Country.csv")
  knows, loc, part = Relation("Knows", 2), Relation("IsLocatedIn", 2), Relation("IsPartOf", 2)
  tri = Relation("Triangle", 3, print_size=True)
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


def q7_program() -> Program:
  """LSQB Q7, the four OPTIONAL-MATCH cases as count-only rules, two with
  negated marker relations (examples/lsqb_q7_optional.py)."""
  C, CM, L, M_, T = (Var(n) for n in ("c", "comment", "liker", "m", "t"))
  has_tag = Relation("HasTag", 2, input_file="Message_hasTag_Tag.csv")
  has_creator = Relation("HasCreator", 2, input_file="Message_hasCreator_Person.csv")
  reply_of = Relation("ReplyOf", 2, input_file="Comment_replyOf_Message.csv")
  likes = Relation("Likes", 2, input_file="Person_likes_Message.csv")
  has_reply, has_liker = Relation("HasReply", 1), Relation("HasLiker", 1)
  c1, c2 = Relation("Case1", 5), Relation("Case2", 4)
  c3, c4 = Relation("Case3", 4), Relation("Case4", 3)
  return Program(rules=[
    (has_reply(M_) <= reply_of(Var("_"), M_)).named("MarkHasReply"),
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


def reach_multi_program() -> Program:
  """Many reachability queries in one fixpoint, the query id as column 0:
  ReachQ(q,y) <= SrcQ(q,y); ReachQ(q,z) <= ReachQ(q,y), Edge(y,z).  Queries
  never mix (q is carried unchanged), so ReachQ restricted to q is the
  single-source result of q's sources."""
  q, y, z = Var("q"), Var("y"), Var("z")
  srcq = Relation("SrcQ", 2, input_file="SrcQ.csv")
  edge = Relation("Edge", 2, input_file="Edge.csv")
  reachq = Relation("ReachQ", 2)
  return Program(rules=[
    (reachq(q, y) <= srcq(q, y)).named("ReachQBase"),
    (reachq(q, z) <= reachq(q, y) & edge(y, z)).named("ReachQRec")
    .with_plan(var_order=["y", "q", "z"]),
  ])


def split_program(split: bool) -> Program:
  """Reachability that skips blocked vertices, with the negation pushed
  below a split (the reference's tests/fixtures/split_rec.nim):
  R(x,y) <= Seed(x,y); R(x,z) <= R(x,y), !Filter(y), [split,] Edge(y,z).
  Without the split marker the same rule compiles to one pipeline; both
  must give the same relation and deltas."""
  x, y, z = Var("x"), Var("y"), Var("z")
  seed = Relation("Seed", 2, input_file="Seed.csv")
  filt = Relation("Filter", 1, input_file="Filter.csv")
  edge = Relation("Edge", 2, input_file="Edge.csv")
  r = Relation("R", 2, print_size=True)
  body = r(x, y) & ~filt(y) & SPLIT & edge(y, z) if split else r(x, y) & ~filt(y) & edge(y, z)
  return Program(rules=[(r(x, y) <= seed(x, y)).named("RBase"), (r(x, z) <= body).named("RRec")])


def cspa_program() -> Program:
  """Context-sensitive points-to analysis (examples/cspa.py; the rules and
  plans of its CSPA_DB program): ValueFlow, ValueAlias and MemoryAlias are
  mutually recursive, one strongly connected component of three relations,
  with wildcard-reflexive base rules."""
  w, x, y, z = Var("w"), Var("x"), Var("y"), Var("z")
  assign_in = Relation("AssignInput", 2, input_file="Assign.csv")
  deref_in = Relation("DerefInput", 2, input_file="Dereference.csv")
  assign = Relation("Assign", 2)
  deref = Relation("Deref", 2)
  vf = Relation("ValueFlow", 2, print_size=True)
  va = Relation("ValueAlias", 2, print_size=True)
  ma = Relation("MemoryAlias", 2, print_size=True)
  return Program(rules=[
    (assign(x, y) <= assign_in(x, y)).named("LoadAssign"),
    (deref(x, y) <= deref_in(x, y)).named("LoadDeref"),
    (vf(y, x) <= assign(y, x)).named("VF_Assign"),
    (vf(x, x) <= assign(x, Var("_"))).named("VF_Refl1"),
    (vf(x, x) <= assign(Var("_"), x)).named("VF_Refl2"),
    (ma(x, x) <= assign(Var("_"), x)).named("MA_Refl1"),
    (ma(x, x) <= assign(x, Var("_"))).named("MA_Refl2"),
    (vf(x, y) <= vf(x, z) & vf(z, y)).named("VF_Trans").with_plan(var_order=["z", "x", "y"]),
    (vf(x, y) <= assign(x, z) & ma(z, y)).named("VF_Assign_MA").with_plan(var_order=["z", "x", "y"]),
    (va(x, y) <= vf(z, x) & vf(z, y)).named("VA_VF").with_plan(var_order=["z", "x", "y"]),
    (va(x, y) <= vf(z, x) & ma(z, w) & vf(w, y)).named("VA_VF_MA_VF")
    .with_plan(var_order=["z", "w", "x", "y"]),
    (ma(x, w) <= deref(y, x) & va(y, z) & deref(z, w)).named("MA_Deref_VA")
    .with_plan(var_order=["y", "z", "x", "w"]),
  ])


def galen_program() -> Program:
  """Galen ontology reasoning (examples/galen.py; the rules and plans of its
  GalenDB program): OutP (pairs) and OutQ (triples) mutually recursive
  through three-atom joins, some with plans for a given DELTA position."""
  e, o, q, r, u, w, x, y, z = (Var(n) for n in ("e", "o", "q", "r", "u", "w", "x", "y", "z"))
  p_in = Relation("PInput", 2, input_file="P.csv")
  q_in = Relation("QInput", 3, input_file="Q.csv")
  r_in = Relation("RInput", 3, input_file="R.csv")
  c_in = Relation("CInput", 3, input_file="C.csv")
  u_in = Relation("UInput", 3, input_file="U.csv")
  s_in = Relation("SInput", 2, input_file="S.csv")
  out_p = Relation("OutP", 2, print_size=True)
  out_q = Relation("OutQ", 3, print_size=True)
  return Program(rules=[
    (out_p(x, z) <= p_in(x, z)).named("LoadP"),
    (out_q(x, r, z) <= q_in(x, r, z)).named("LoadQ"),
    (out_p(x, z) <= out_p(x, y) & out_p(y, z)).named("TC"),
    (out_q(x, r, z) <= out_p(x, y) & out_q(y, r, z)).named("PropQ"),
    (out_p(x, z) <= out_p(y, w) & u_in(w, r, z) & out_q(x, r, y)).named("Join3a")
    .with_plan(delta=2, var_order=["y", "w", "r", "x", "z"]),
    (out_p(x, z) <= c_in(y, w, z) & out_p(x, w) & out_p(x, y)).named("Join3b"),
    (out_q(x, q, z) <= out_q(x, r, z) & s_in(r, q)).named("PropQS"),
    (out_q(x, e, o) <= out_q(x, y, z) & r_in(y, u, e) & out_q(z, u, o)).named("Join3c")
    .with_plan(delta=0, var_order=["z", "y", "u", "x", "e", "o"]),
  ])


# Constants of typelookup_program (ids a fact extractor would assign).
TL_OBJECT, TL_ABSTRACT, TL_STATIC, TL_CLINIT, TL_MAIN = 0, 1, 2, 3, 7


def typelookup_program() -> Program:
  """Subtyping, virtual-method lookup and class initialisation in the
  style of DOOP's type rules (the reference's examples/doop.py), written for
  this package's tests: constants in rule heads (every class is a subtype of
  the root type), in body atoms (the main method's name and its static
  modifier) and in negations (methods not declared abstract), a negation
  with a wildcard (no method of that name declared in the class), a
  repeated head variable (reflexive subtyping), a filter, and three
  recursive relations, one of them over a stratified negation."""
  s, t, u, m, n = Var("s"), Var("t"), Var("u"), Var("m"), Var("n")
  sup = Relation("DirectSuperclass", 2, input_file="DirectSuperclass.csv")
  cls = Relation("ClassType", 1, input_file="ClassType.csv")
  decl = Relation("MethodDecl", 2, input_file="MethodDecl.csv")          # (method, class)
  name = Relation("MethodName", 2, input_file="MethodName.csv")          # (method, name)
  mod = Relation("MethodModifier", 2, input_file="MethodModifier.csv")   # (modifier, method)
  main_cls = Relation("MainClass", 1, input_file="MainClass.csv")
  sub = Relation("SubtypeOf", 2, print_size=True)
  impl = Relation("Implemented", 3, print_size=True)                     # (class, name, method)
  look = Relation("Lookup", 3, print_size=True)
  main = Relation("MainMethod", 1, print_size=True)
  init = Relation("Initialized", 1, print_size=True)
  return Program(rules=[
    (sub(s, s) <= cls(s)).named("SubRefl"),
    (sub(s, t) <= sup(s, t)).named("SubDirect"),
    (sub(s, t) <= sub(s, u) & sup(u, t)).named("SubTrans"),
    (sub(s, Const(TL_OBJECT)) <= cls(s)).named("SubRoot"),
    (impl(t, n, m) <= decl(m, t) & name(m, n) & ~mod(Const(TL_ABSTRACT), m)).named("Implemented"),
    (look(t, n, m) <= impl(t, n, m)).named("LookupOwn"),
    (look(t, n, m) <= sup(t, s) & look(s, n, m) & ~impl(t, n, Var("_"))).named("LookupInherited"),
    (main(m) <= main_cls(t) & decl(m, t) & name(m, Const(TL_MAIN)) & mod(Const(TL_STATIC), m)
     & Filter(("m",), f"return m != {TL_CLINIT}u;")).named("MainMethod"),
    (init(t) <= main(m) & decl(m, t)).named("InitMain"),
    (init(s) <= init(t) & sup(t, s)).named("InitSuper"),
  ])


def split_nonrec_program() -> Program:
  """Q(x,z) <= A(x,y), !B(y), split, C(x,z) (tests/fixtures/split_test.nim):
  the split of a non-recursive rule."""
  x, y, z = Var("x"), Var("y"), Var("z")
  a = Relation("A", 2, input_file="A.csv")
  b = Relation("B", 1, input_file="B.csv")
  c = Relation("C", 2, input_file="C.csv")
  q = Relation("Q", 2, print_size=True)
  return Program(rules=[(q(x, z) <= a(x, y) & ~b(y) & SPLIT & c(x, z)).named("SplitTest")])


PROGRAMS = {
  "tc": tc_program,
  "sg": sg_program,
  "andersen": andersen_program,
  "triangle": lambda: triangle_program(False),
  "triangle_count": lambda: triangle_program(True),
  "reach": reach_program,
  "reach_multi": reach_multi_program,
  "neg2hop": lambda: neg2hop_program(False),
  "neg2hop_count": lambda: neg2hop_program(True),
  "twohop": lambda: twohop_program(False),
  "twohop_count": lambda: twohop_program(True),
  "twohop_p2": lambda: twohop_program(False, p2_first=True),
  "q7": q7_program,
  "q3": q3_program,
  "split_rec": lambda: split_program(True),
  "split_rec_plain": lambda: split_program(False),
  "split_nonrec": split_nonrec_program,
  "typelookup": typelookup_program,
  "cspa": cspa_program,
  "galen": galen_program,
}


def so_path(name: str) -> str:
  return os.path.join(PROG_DIR, f"lib{name}.so")


def build_one(name: str) -> str:
  prog = PROGRAMS[name]()
  proj = B.build_project(prog, name, cache_base=os.path.join(PROG_DIR, "src"))
  res = B.compile_jit_project(proj, out_path=so_path(name))
  if not res.ok():
    raise RuntimeError(f"compiling program {name} failed:\n{res.stderr[-6000:]}")
  return res.artifact


def build_all(jobs: int = 6) -> list[str]:
  os.makedirs(PROG_DIR, exist_ok=True)
  src_dir = os.path.join(PROG_DIR, "src")
  before = set(os.listdir(src_dir)) if os.path.isdir(src_dir) else set()
  with ThreadPoolExecutor(max_workers=jobs) as ex:
    out = list(ex.map(build_one, PROGRAMS))
  # generated sources of earlier builds (named by content hash) are stale
  for d in before - set(os.path.basename(os.path.dirname(p)) for p in _latest_sources(src_dir)):
    shutil.rmtree(os.path.join(src_dir, d), ignore_errors=True)
  return out


def _latest_sources(src_dir: str) -> list[str]:
  """The newest generated source file of each bundled program."""
  out = []
  for name in PROGRAMS:
    cands = [os.path.join(src_dir, d) for d in os.listdir(src_dir)
             if d.rsplit("_", 1)[0] == name]
    if cands:
      newest = max(cands, key=os.path.getmtime)
      out += [os.path.join(newest, f) for f in os.listdir(newest)][:1]
  return out


_loaded: dict[str, ProgramLib] = {}


def load(name: str) -> ProgramLib:
  """ctypes binding of a prebuilt program library (build_all() first)."""
  if name not in _loaded:
    path = so_path(name)
    if not os.path.exists(path):
      raise RuntimeError(f"{path} is not built; run __graft_entry__.build() first")
    _loaded[name] = ProgramLib(path)
  return _loaded[name]
