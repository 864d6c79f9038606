"""Pure-Python naive Jacobi evaluation for tiny inputs.  Test infrastructure only.

Each program is its rule set written as Python set comprehensions.  The
recursive step iterates F_{i+1} = F_i u T(F_i) from the base-stratum result
F_0 and records |T(F_i) - F_i| per iteration, stopping after the first empty
difference -- the delta sequence the reference's semi-naive loop produces
(iteration 0 reads FULL for DELTA, every other atom reads FULL; see the
oracle.cpp header for the reference citations).  It shares no code with the
C++ oracle, so agreement of the two pins the restatement.
"""

from __future__ import annotations


def _fix(full, step, max_iters=None):
  deltas = []
  it = 0
  while max_iters is None or it < max_iters:
    new = step(full) - full
    deltas.append(len(new))
    full = full | new
    it += 1
    if not new:
      break
  return full, deltas


def tc(arc):
  edge = {(int(a), int(b)) for a, b in arc}
  succ = {}
  for a, b in edge:
    succ.setdefault(a, set()).add(b)
  path0 = set(edge)
  path, deltas = _fix(path0, lambda p: {(x, z) for (x, y) in p for z in succ.get(y, ())})
  return {"Edge": edge, "Path": path}, {"Path": deltas}


def sg(arc):
  arc = {(int(a), int(b)) for a, b in arc}
  kids = {}
  for p, x in arc:
    kids.setdefault(p, set()).add(x)
  base = {(x, y) for p, ks in kids.items() for x in ks for y in ks if x != y}
  sgr, deltas = _fix(base, lambda s: {(x, y) for (a, b) in s
                                     for x in kids.get(a, ()) for y in kids.get(b, ())})
  return {"Sg": sgr}, {"Sg": deltas}


def andersen(addr, assign, load, store):
  addr = {(int(a), int(b)) for a, b in addr}
  assign = {(int(a), int(b)) for a, b in assign}
  load = {(int(a), int(b)) for a, b in load}
  store = {(int(a), int(b)) for a, b in store}

  def step(pt):
    by_v = {}
    for v, o in pt:
      by_v.setdefault(v, set()).add(o)
    out = {(v, o) for (v, w) in assign for o in by_v.get(w, ())}
    out |= {(x, o) for (x, y) in load for a in by_v.get(y, ()) for o in by_v.get(a, ())}
    out |= {(a, o) for (x, y) in store for a in by_v.get(x, ()) for o in by_v.get(y, ())}
    return out

  pt, deltas = _fix(set(addr), step)
  return {"PointsTo": pt}, {"PointsTo": deltas}


def triangle(rows):
  kn = {(int(a), int(b)) for a, b in rows}
  succ = {}
  for a, b in kn:
    succ.setdefault(a, set()).add(b)
  tri = {(a, b, c) for (a, b) in kn for c in succ.get(b, ()) if c in succ.get(a, ())}
  return {"Knows": kn, "Triangle": tri}, {"Triangle": [len(tri)]}


def reach(edge, src):
  edge = {(int(a), int(b)) for a, b in edge}
  succ = {}
  for a, b in edge:
    succ.setdefault(a, set()).add(b)
  r0 = {int(s) for s in src}
  r, deltas = _fix(r0, lambda r: {z for y in r for z in succ.get(y, ())})
  return {"Reach": r}, {"Reach": deltas}


def neg2hop(knows, interest):
  """LSQB Q9: Path(p1,p2,p3,t) <= Knows(p1,p2), Knows(p2,p3), HasInterest(p3,t),
  not Knows(p1,p3), p1 != p3."""
  kn = {(int(a), int(b)) for a, b in knows}
  hi = {(int(a), int(b)) for a, b in interest}
  succ, tags = {}, {}
  for a, b in kn:
    succ.setdefault(a, set()).add(b)
  for p, t in hi:
    tags.setdefault(p, set()).add(t)
  path = {(p1, p2, p3, t) for (p1, p2) in kn for p3 in succ.get(p2, ())
          if p1 != p3 and (p1, p3) not in kn for t in tags.get(p3, ())}
  return {"Knows": kn, "HasInterest": hi, "Path": path}, {"Path": [len(path)]}


def twohop(knows, interest):
  """LSQB Q6 (examples/lsqb_q6_2hop.py): Path(p1,p2,p3,t) <= Knows(p1,p2),
  Knows(p2,p3), HasInterest(p3,t), p1 != p3."""
  kn = {(int(a), int(b)) for a, b in knows}
  hi = {(int(a), int(b)) for a, b in interest}
  succ, tags = {}, {}
  for a, b in kn:
    succ.setdefault(a, set()).add(b)
  for p, t in hi:
    tags.setdefault(p, set()).add(t)
  path = {(p1, p2, p3, t) for (p1, p2) in kn for p3 in succ.get(p2, ())
          if p1 != p3 for t in tags.get(p3, ())}
  return {"Knows": kn, "HasInterest": hi, "Path": path}, {"Path": [len(path)]}


def twohop_as_planned(knows, interest):
  """What the plan order p1, p2, p3, t compiles LSQB Q6 to (the reference's
  tests/fixtures/integration/lsqb_q6_2hop.mir.sexpr): p2 is joined on column
  0 of Knows' (0 1) index and p1 drawn from the same index under prefix p2,
  so Path(p1,p2,p3,t) <= Knows(p2,p1), Knows(p2,p3), HasInterest(p3,t),
  p1 != p3."""
  kn = {(int(a), int(b)) for a, b in knows}
  succ, tags = {}, {}
  for a, b in kn:
    succ.setdefault(a, set()).add(b)
  for p, t in interest:
    tags.setdefault(int(p), set()).add(int(t))
  return {(p1, p2, p3, t) for p2, s in succ.items() for p1 in s for p3 in s if p1 != p3
          for t in tags.get(p3, ())}


def twohop_as_planned_count_closed_form(knows, interest, n):
  """|twohop_as_planned|: for each p2, every ordered pair of distinct
  successors (p1, p3) times the tags of p3 = sum_p (deg+(p) - 1) (A t)[p]."""
  import numpy as np
  e = np.unique(np.asarray(knows, dtype=np.int64), axis=0)
  h = np.unique(np.asarray(interest, dtype=np.int64), axis=0)
  t = np.bincount(h[:, 0], minlength=n).astype(np.int64)
  deg = np.bincount(e[:, 0], minlength=n).astype(np.int64)
  at = np.bincount(e[:, 0], weights=t[e[:, 1]], minlength=n).astype(np.int64)
  return int(((deg - 1) * at).sum())


def twohop_count_closed_form(knows, interest, n):
  """|Path| of LSQB Q6 without enumeration: with A the n x n 0/1 Knows
  matrix and t[p] the number of tags of p, the sum over p1 != p3 of
  (A A)[p1, p3] t[p3] = 1' A A t - sum_p (A A)[p, p] t[p] (sparse products)."""
  import numpy as np
  import scipy.sparse as sp
  e = np.unique(np.asarray(knows, dtype=np.int64), axis=0)
  a = sp.csr_matrix((np.ones(e.shape[0], np.int64), (e[:, 0], e[:, 1])), shape=(n, n))
  h = np.unique(np.asarray(interest, dtype=np.int64), axis=0)
  t = np.bincount(h[:, 0], minlength=n).astype(np.int64)
  total = int((a @ (a @ t)).sum())
  diag = np.asarray((a.multiply(a.T)).sum(axis=1)).ravel()   # (A A)[p, p] = sum_q A[p,q] A[q,p]
  return total - int((diag * t).sum())


def q7(has_tag, has_creator, reply_of, likes):
  """LSQB Q7 by enumerating each case's distinct head tuples."""
  ht = {(int(a), int(b)) for a, b in has_tag}
  hc = {(int(a), int(b)) for a, b in has_creator}
  ro = {(int(a), int(b)) for a, b in reply_of}      # (comment, m)
  lk = {(int(a), int(b)) for a, b in likes}         # (liker, m)
  has_reply = {m for _, m in ro}
  has_liker = {m for _, m in lk}
  c1 = {(m, t, c, cm, l) for (m, t) in ht for (m2, c) in hc if m2 == m
        for (cm, m3) in ro if m3 == m for (l, m4) in lk if m4 == m}
  c2 = {(m, t, c, cm) for (m, t) in ht for (m2, c) in hc if m2 == m
        for (cm, m3) in ro if m3 == m if m not in has_liker}
  c3 = {(m, t, c, l) for (m, t) in ht for (m2, c) in hc if m2 == m
        for (l, m4) in lk if m4 == m if m not in has_reply}
  c4 = {(m, t, c) for (m, t) in ht for (m2, c) in hc if m2 == m
        if m not in has_reply and m not in has_liker}
  return {"Case1": len(c1), "Case2": len(c2), "Case3": len(c3), "Case4": len(c4)}


def q3(knows, located, part_of):
  """LSQB Q3: directed 3-cycles p1->p2->p3->p1 whose persons' cities lie in
  one common country."""
  kn = {(int(a), int(b)) for a, b in knows}
  loc = {(int(a), int(b)) for a, b in located}
  part = {(int(a), int(b)) for a, b in part_of}
  tri = {(p1, p2, p3) for (p1, p2) in kn for (q, p3) in kn if q == p2 and (p3, p1) in kn
         for (x1, c1) in loc if x1 == p1 for (x2, c2) in loc if x2 == p2
         for (x3, c3) in loc if x3 == p3 for (d1, co) in part if d1 == c1
         if (c2, co) in part and (c3, co) in part}
  return {"Triangle": tri}


def split_rec(seed, filt, edge):
  """R(x,y) <= Seed(x,y); R(x,z) <= R(x,y), !Filter(y), Edge(y,z)."""
  seed = {(int(a), int(b)) for a, b in seed}
  blocked = {int(v[0]) if hasattr(v, "__len__") else int(v) for v in filt}
  succ = {}
  for a, b in edge:
    succ.setdefault(int(a), set()).add(int(b))
  r, deltas = _fix(set(seed), lambda p: {(x, z) for (x, y) in p if y not in blocked
                                         for z in succ.get(y, ())})
  return {"R": r}, {"R": deltas}


def split_nonrec(a, b, c):
  """Q(x,z) <= A(x,y), !B(y), C(x,z): one pass, no iteration."""
  blocked = {int(v[0]) if hasattr(v, "__len__") else int(v) for v in b}
  keep = {int(x) for x, y in a if int(y) not in blocked}
  q = {(int(x), int(z)) for x, z in c if int(x) in keep}
  return {"Q": q}, {}


def typelookup(sup, cls, decl, name, mod, main_cls, obj=0, abstract=1, static=2, clinit=3,
               main_name=7):
  """The type-lookup program of the package's tests (subtyping, method
  lookup, class initialisation in DOOP's style) as set comprehensions under
  stratified semantics: each recursive relation is iterated to its least
  fixpoint after the relations it negates are complete.  Constants restrict
  body columns, negated constants are part of the excluded prefix, head
  constants are emitted as values (semantics of the reference's Const and
  negation, src/srdatalog/dsl.py:63-100, hir/plan.py:251-285)."""
  sup = {(int(a), int(b)) for a, b in sup}
  cls = {int(c) for (c,) in cls}
  decl = {(int(a), int(b)) for a, b in decl}
  name = {(int(a), int(b)) for a, b in name}
  mod = {(int(a), int(b)) for a, b in mod}
  main_cls = {int(c) for (c,) in main_cls}

  def closure(base, step):
    full = set(base)
    while True:
      new = step(full) - full
      if not new:
        return full
      full |= new

  up = {}
  for a, b in sup:
    up.setdefault(a, set()).add(b)
  sub = closure({(c, c) for c in cls} | sup | {(c, obj) for c in cls},
                lambda f: {(s, t) for s, u in f for t in up.get(u, ())})
  impl = {(t, n, m) for m, t in decl for m2, n in name if m2 == m and (abstract, m) not in mod}
  own = {(t, n) for t, n, _ in impl}
  look = closure(impl, lambda f: {(t, n, m) for t, s in sup for s2, n, m in f
                                  if s2 == s and (t, n) not in own})
  main = {m for t in main_cls for m, t2 in decl if t2 == t and (m, main_name) in name
          and (static, m) in mod and m != clinit}
  cls_of = {}
  for m, t in decl:
    cls_of.setdefault(m, set()).add(t)
  init = closure({t for m in main for t in cls_of.get(m, ())},
                 lambda f: {s for t in f for s in up.get(t, ())})
  return {"SubtypeOf": sub, "Implemented": impl, "Lookup": look, "MainMethod": {(m,) for m in main},
          "Initialized": {(t,) for t in init}}


def cspa(assign, deref):
  """CSPA (the reference's examples/cspa.py): the stratum's base rules give
  F_0, then the three mutually recursive relations are iterated together,
  F_{i+1} = F_i u T(F_i) with every rule reading F_i, recording
  |T(F_i) - F_i| per relation and iteration until all three are empty --
  the delta sequences of the reference's semi-naive loop over one strongly
  connected component (iteration 0 reads FULL for DELTA)."""
  a = {(int(p), int(q)) for p, q in assign}
  d = {(int(p), int(q)) for p, q in deref}
  vf = set(a) | {(x, x) for x, _ in a} | {(x, x) for _, x in a}
  ma = {(x, x) for _, x in a} | {(x, x) for x, _ in a}
  va: set = set()

  def by_first(rel):
    out: dict = {}
    for p, q in rel:
      out.setdefault(p, set()).add(q)
    return out

  deltas = {"ValueFlow": [], "ValueAlias": [], "MemoryAlias": []}
  a1, d1 = by_first(a), by_first(d)
  while True:
    vf1, ma1, va1 = by_first(vf), by_first(ma), by_first(va)
    t_vf = {(x, y) for x, z in vf for y in vf1.get(z, ())}
    t_vf |= {(x, y) for x, z in a for y in ma1.get(z, ())}
    t_va = {(x, y) for z, xs in vf1.items() for x in xs for y in xs}
    t_va |= {(x, y) for z, xs in vf1.items() for w in ma1.get(z, ()) for x in xs
             for y in vf1.get(w, ())}
    t_ma = {(x, w) for y, xs in d1.items() for z in va1.get(y, ()) for x in xs
            for w in d1.get(z, ())}
    new = (t_vf - vf, t_va - va, t_ma - ma)
    deltas["ValueFlow"].append(len(new[0]))
    deltas["ValueAlias"].append(len(new[1]))
    deltas["MemoryAlias"].append(len(new[2]))
    vf, va, ma = vf | new[0], va | new[1], ma | new[2]
    if not any(new):
      break
  return {"ValueFlow": vf, "ValueAlias": va, "MemoryAlias": ma}, deltas


def galen(p_in, q_in, r_in, c_in, u_in, s_in):
  """Galen (the reference's examples/galen.py): F_0 = the loaded P and Q,
  then OutP and OutQ iterated together, F_{i+1} = F_i u T(F_i), recording
  |T(F_i) - F_i| per relation and iteration until both are empty."""
  P = {(int(a), int(b)) for a, b in p_in}
  Q = {(int(a), int(b), int(c)) for a, b, c in q_in}
  R = [(int(a), int(b), int(c)) for a, b, c in r_in]
  C = [(int(a), int(b), int(c)) for a, b, c in c_in]
  U = [(int(a), int(b), int(c)) for a, b, c in u_in]
  S = [(int(a), int(b)) for a, b in s_in]
  deltas = {"OutP": [], "OutQ": []}
  while True:
    p_from, p_to = {}, {}
    for a, b in P:
      p_from.setdefault(a, set()).add(b)
      p_to.setdefault(b, set()).add(a)
    q_from = {}
    q_by_x_r = {}
    for a, b, c in Q:
      q_from.setdefault(a, set()).add((b, c))
      q_by_x_r.setdefault((a, b), set()).add(c)
    q_by_r_y: dict = {}
    for x, r, y in Q:
      q_by_r_y.setdefault((r, y), set()).add(x)
    t_p = {(x, z) for x, y in P for z in p_from.get(y, ())}                       # TC
    t_p |= {(x, z) for w, r, z in U for y in p_to.get(w, ()) for x in q_by_r_y.get((r, y), ())}  # Join3a
    t_p |= {(x, z) for y, w, z in C for x in p_to.get(w, ()) if (x, y) in P}      # Join3b
    t_q = {(x, r, z) for x, y in P for r, z in q_from.get(y, ())}                 # PropQ
    s_of: dict = {}
    for r, q in S:
      s_of.setdefault(r, set()).add(q)
    t_q |= {(x, q, z) for x, r, z in Q for q in s_of.get(r, ())}                  # PropQS
    r_of: dict = {}
    for y, u, e in R:
      r_of.setdefault(y, []).append((u, e))
    t_q |= {(x, e, o) for x, y, z in Q for u, e in r_of.get(y, ())                 # Join3c
            for o in q_by_x_r.get((z, u), ())}
    new_p, new_q = t_p - P, t_q - Q
    deltas["OutP"].append(len(new_p))
    deltas["OutQ"].append(len(new_q))
    P |= new_p
    Q |= new_q
    if not new_p and not new_q:
      break
  return {"OutP": P, "OutQ": Q}, deltas
