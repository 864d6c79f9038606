"""DSL Program -> MIR, following the reference's planning semantics.

The plan this produces is the contract the backend executes, so it must be
the plan the reference would run: same strata, semi-naive variants, join
order, column indexes and maintenance op order.  The stages restate the
reference pipeline (src/srdatalog/hir/__init__.py compile_to_mir):

  stratify         Tarjan SCCs over head->body IDB edges, visited in rule
                   order with sorted neighbours; a recursive SCC's base
                   rules become their own non-recursive stratum first;
                   consecutive independent non-recursive strata fuse
                   (hir/stratify.py:225-254)
  variants         base variant per non-recursive rule; one DELTA variant
                   per SCC atom of a recursive rule (hir/semi_naive.py:47)
  planning         clause order (filters first, then the delta atom, then
                   max overlap with bound vars), var order (delta atom's
                   join vars, other join vars, free vars), user plans,
                   per-atom access order / index columns (hir/plan.py)
  index selection  required indexes per SCC member and a canonical index,
                   preferring one read as FULL (hir/index.py:104)
  lowering         column-join per join var, one cartesian-join for the
                   free vars, filters, insert-into; maintenance op order
                   (hir/lower.py:487-926)
  MIR passes       rebuilds before reconstruct, source reorder by clause
                   order then prefixed-first (mir/passes.py:104-340)

The result prints byte-identically to the reference's MIR goldens
(tests/test_compiler_golden.py).
"""

from __future__ import annotations

from dataclasses import dataclass, field

from . import mir as M
from .codegen import RelDecl
from .dsl import Atom, Filter, Negation, Program, Rule, Split, const_value, is_const


class CompileError(ValueError):
  pass


# -------------------------------------------------------------- strata ----
@dataclass
class Stratum:
  members: set[str]
  recursive: bool
  rules: list[Rule]
  base_variants: list["Variant"] = field(default_factory=list)
  rec_variants: list["Variant"] = field(default_factory=list)
  required: dict[str, list[list[int]]] = field(default_factory=dict)
  canonical: dict[str, list[int]] = field(default_factory=dict)


def _body_rels(rule: Rule) -> set[str]:
  out = set()
  for c in rule.body:
    if isinstance(c, Atom):
      out.add(c.rel)
    elif isinstance(c, Negation):
      out.add(c.atom.rel)
  return out


def stratify(rules: list[Rule]) -> list[Stratum]:
  idbs = {h.rel for r in rules for h in r.heads}
  deps: dict[str, set[str]] = {rel: set() for rel in idbs}
  for r in rules:
    b = _body_rels(r) & idbs
    for h in r.heads:
      deps[h.rel] |= b

  # Tarjan, iterative over an explicit stack; emits SCCs bottom-up.
  order, low, on, stack, sccs = {}, {}, set(), [], []
  counter = 0

  def visit(root: str) -> None:
    nonlocal counter
    work = [(root, iter(sorted(deps[root])))]
    order[root] = low[root] = counter
    counter += 1
    stack.append(root)
    on.add(root)
    while work:
      v, it = work[-1]
      advanced = False
      for w in it:
        if w not in order:
          order[w] = low[w] = counter
          counter += 1
          stack.append(w)
          on.add(w)
          work.append((w, iter(sorted(deps[w]))))
          advanced = True
          break
        if w in on:
          low[v] = min(low[v], order[w])
      if advanced:
        continue
      work.pop()
      if work:
        u = work[-1][0]
        low[u] = min(low[u], low[v])
      if low[v] == order[v]:
        comp = set()
        while True:
          w = stack.pop()
          on.discard(w)
          comp.add(w)
          if w == v:
            break
        sccs.append(comp)

  for r in rules:
    for h in r.heads:
      if h.rel not in order:
        visit(h.rel)

  strata: list[Stratum] = []
  for comp in sccs:
    comp_rules = [r for r in rules if any(h.rel in comp for h in r.heads)]
    if not comp_rules:
      continue
    rec = len(comp) > 1 or any(rel in deps[rel] for rel in comp)
    if rec:
      base = [r for r in comp_rules if not (_body_rels(r) & comp)]
      recr = [r for r in comp_rules if _body_rels(r) & comp]
      if base:
        strata.append(Stratum(set(comp), False, base))
      if recr:
        strata.append(Stratum(set(comp), True, recr))
    else:
      strata.append(Stratum(set(comp), False, comp_rules))
  return _fuse(strata)


def _fuse(strata: list[Stratum]) -> list[Stratum]:
  rec_members = set().union(*[s.members for s in strata if s.recursive]) if strata else set()

  def fusable(s):
    return not s.recursive and not (s.members & rec_members)

  out, i = [], 0
  while i < len(strata):
    cur = strata[i]
    if not fusable(cur):
      out.append(cur)
      i += 1
      continue
    produced = set(cur.members)
    j = i + 1
    while j < len(strata) and fusable(strata[j]):
      nxt = strata[j]
      if any(_body_rels(r) & produced for r in nxt.rules):
        break
      cur.members |= nxt.members
      cur.rules.extend(nxt.rules)
      produced = set(cur.members)
      j += 1
    out.append(cur)
    i = j
  return out


# ------------------------------------------------------------ planning ----
@dataclass
class Access:
  rel: str
  ver: str
  order: list[str]        # variables in access order
  index: list[int]        # column permutation
  prefix_len: int
  clause: int


@dataclass
class Variant:
  rule: Rule
  delta: int
  versions: list[str]
  clause_order: list[int] = field(default_factory=list)
  var_order: list[str] = field(default_factory=list)
  join_vars: set[str] = field(default_factory=set)
  atoms: list[Access] = field(default_factory=list)
  negs: list[Access] = field(default_factory=list)
  split_at: int = -1          # body index of the SPLIT marker (-1: none)
  temp_vars: list[str] = field(default_factory=list)
  temp_rel: str = ""


def _clause_vars(c) -> list[str]:
  if isinstance(c, Atom):
    return [v for v in c.args if not is_const(v)]
  if isinstance(c, Negation):
    return [v for v in c.atom.args if not is_const(v)]
  if isinstance(c, Filter):
    return list(c.vars)
  return []


def _needs(c) -> set[str]:
  if isinstance(c, (Negation, Filter)):
    return set(_clause_vars(c))
  return set()


def _binds(c) -> set[str]:
  return set(_clause_vars(c)) if isinstance(c, Atom) else set()


def _join_vars(rule: Rule) -> set[str]:
  seen: dict[str, int] = {}
  for c in rule.body:
    if isinstance(c, Atom):
      for v in _clause_vars(c):
        seen[v] = seen.get(v, 0) + 1
  return {v for v, n in seen.items() if n > 1}


def _heuristic_clause_order(rule: Rule, delta: int) -> list[int]:
  bound: set[str] = set()
  left = set(range(len(rule.body)))
  out = []
  while left:
    ready = sorted(i for i in left if _needs(rule.body[i]) <= bound)
    if not ready:
      atoms = [i for i in sorted(left) if isinstance(rule.body[i], Atom)]
      ready = [atoms[0] if atoms else min(left)]
    pick = next((i for i in ready if isinstance(rule.body[i], Filter)), -1)
    if pick < 0 and delta in ready:
      pick = delta
    if pick < 0:
      best = -1
      for i in ready:
        c = rule.body[i]
        ov = sum(1 for v in _clause_vars(c) if v in bound) if isinstance(c, Atom) else 0
        if ov > best:
          best, pick = ov, i
    out.append(pick)
    left.discard(pick)
    bound |= _binds(rule.body[pick])
  return out


def _var_order(rule: Rule, clause_order: list[int], jv: set[str], delta: int) -> list[str]:
  out: list[str] = []

  def add(vs, want_join: bool):
    for v in vs:
      if (v in jv) == want_join and v not in out:
        out.append(v)

  if 0 <= delta < len(rule.body):
    add(_clause_vars(rule.body[delta]), True)
  for i in clause_order:
    add(_clause_vars(rule.body[i]), True)
  for i in clause_order:
    add(_clause_vars(rule.body[i]), False)
  return out


def _clause_order_from_vars(rule: Rule, var_order: list[str], delta: int) -> list[int]:
  out, left, bound = [], set(range(len(rule.body))), set()
  for v in var_order:
    if v in bound:
      continue
    cands = sorted(i for i in left if _needs(rule.body[i]) <= bound
                   and v in _clause_vars(rule.body[i]))
    if not cands:
      continue
    pick = delta if delta in cands else cands[0]
    out.append(pick)
    left.discard(pick)
    bound |= set(_clause_vars(rule.body[pick]))
  while left:
    ready = [i for i in sorted(left) if _needs(rule.body[i]) <= bound]
    filt = [i for i in ready if isinstance(rule.body[i], Filter)]
    pick = filt[0] if filt else (ready[0] if ready else min(left))
    out.append(pick)
    left.discard(pick)
    bound |= set(_clause_vars(rule.body[pick]))
  return out


def _access(atom: Atom, ver: str, jv: set[str], var_order: list[str], clause: int,
            negated: bool) -> Access:
  cvars = {v for v in atom.args if not is_const(v)}
  order = [v for v in var_order if v in cvars] + sorted(cvars - set(var_order))
  index = [atom.args.index(v) for v in order]
  if negated:
    prefix_len = sum(1 for v in order if not v.startswith("_"))
    # constants (kept in negated atoms) lead the index and the prefix
    # (reference: hir/plan.py:283-285)
    consts = [c for c, v in enumerate(atom.args) if is_const(v)]
    order = [atom.args[c] for c in consts] + order
    index = consts + index
    prefix_len += len(consts)
  else:
    if any(is_const(v) for v in atom.args):
      raise CompileError(f"{atom.rel}: a constant in a positive atom was not bound")
    prefix_len = sum(1 for v in order if v in jv)
  index += [c for c in range(len(atom.args)) if c not in index]
  return Access(atom.rel, ver, order, index, prefix_len, clause)


def plan_variant(v: Variant) -> None:
  rule = v.rule
  jv = _join_vars(rule)
  plan = next((p for p in rule.plans if p.delta == v.delta), None)
  if plan is not None and plan.var_order:
    v.var_order = list(plan.var_order)
    v.clause_order = (list(plan.clause_order) if plan.clause_order
                      else _clause_order_from_vars(rule, v.var_order, v.delta))
  else:
    v.clause_order = _heuristic_clause_order(rule, v.delta)
    v.var_order = _var_order(rule, v.clause_order, jv, v.delta)
  v.join_vars = jv
  v.split_at = detect_split(rule)
  if v.split_at >= 0 and rule.name:
    v.temp_vars = split_temp_vars(rule, v.split_at)
    v.temp_rel = f"_temp_{rule.name}"
  for k, c in enumerate(rule.body):
    if isinstance(c, Atom):
      v.atoms.append(_access(c, v.versions[k], jv, v.var_order, k, False))
    elif isinstance(c, Negation):
      v.negs.append(_access(c.atom, M.FULL, jv, v.var_order, k, True))


# ------------------------------------------------------------- splits ----
def detect_split(rule: Rule) -> int:
  """Body index of the SPLIT marker, -1 if none (reference: hir/plan.py:385)."""
  return next((i for i, c in enumerate(rule.body) if isinstance(c, Split)), -1)


def split_temp_vars(rule: Rule, split_at: int) -> list[str]:
  """Variables crossing the split: bound by an atom or negation above it and
  used below it (by an atom, a negation or a head).  Those the atoms below
  also use come first, then the head-only ones, each group sorted -- the
  column order of the temp relation (reference: hir/plan.py:393-436)."""
  above: set[str] = set()
  for c in rule.body[:split_at]:
    if isinstance(c, (Atom, Negation)):
      above.update(_clause_vars(c))
  below_atoms: set[str] = set()
  for c in rule.body[split_at + 1:]:
    if isinstance(c, Atom):
      below_atoms.update(c.args)
  below = set(below_atoms)
  for h in rule.heads:
    below.update(_clause_vars(h))
  for c in rule.body[split_at + 1:]:
    if isinstance(c, Negation):
      below.update(c.atom.args)
  first = [x for x in sorted(above) if x in below_atoms and x in below]
  return first + [x for x in sorted(above) if x in below and x not in below_atoms]


def split_temps(rules) -> list[tuple[str, int]]:
  """(name, arity) of every split rule's temp relation, in rule order."""
  out = []
  for r in rules:
    at = detect_split(r)
    if at >= 0 and r.name:
      name = f"_temp_{r.name}"
      if all(name != n for n, _ in out):
        out.append((name, len(split_temp_vars(r, at))))
  return out


def _split_above(v: Variant) -> list:
  """First pipeline of a split rule: the one positive atom above the split
  as a scan, the negations and filters above it, written into the temp.
  Empty when more than one atom is above the split (the rule is then
  compiled unsplit), as the reference does (hir/lower.py:288-350)."""
  above = [a for a in v.atoms if a.clause < v.split_at]
  if len(above) != 1:
    return []
  a = above[0]
  ops: list = [M.Scan(list(a.order), a.rel, list(a.index), a.ver, [])]
  for n in v.negs:
    if n.clause < v.split_at:
      ops.append(M.Negation(n.rel, M.FULL, list(n.index), list(n.order[:n.prefix_len])))
  for c in v.rule.body[:v.split_at]:
    if isinstance(c, Filter):
      ops.append(M.Filter(list(c.vars), c.code))
  ops.append(M.InsertInto(v.temp_rel, M.NEW, list(range(len(v.temp_vars))), list(v.temp_vars)))
  return ops


def _split_below(v: Variant, s: Stratum, temp_ver: str) -> list:
  """Second pipeline: scan the temp, then one cartesian level over the
  atoms below the split that bind head variables (each prefixed by the temp
  variables it shares), the negations and filters below, the inserts
  (reference: hir/lower.py:353-433).  Like the reference, an atom below the
  split that binds no head variable is not joined."""
  tv = list(v.temp_vars)
  tset = set(tv)
  ops: list = [M.Scan(tv, v.temp_rel, list(range(len(tv))), temp_ver, [])]
  heads = {x for h in v.rule.heads for x in _clause_vars(h)}
  cvars: list[str] = []
  srcs, vfs = [], []
  for a in v.atoms:
    if a.clause <= v.split_at:
      continue
    mine = [x for x in a.order if x not in tset and x in heads and x not in cvars]
    if not mine:
      continue
    srcs.append(M.ColumnSource(a.rel, list(a.index), a.ver, [x for x in a.order if x in tset],
                               clause_idx=a.clause))
    vfs.append(mine)
    cvars += mine
  if cvars:
    ops.append(M.CartesianJoin(cvars, vfs, srcs))
  for n in v.negs:
    if n.clause > v.split_at:
      ops.append(M.Negation(n.rel, M.FULL, list(n.index), list(n.order[:n.prefix_len])))
  for c in v.rule.body[v.split_at + 1:]:
    if isinstance(c, Filter):
      ops.append(M.Filter(list(c.vars), c.code))
  return ops + _inserts(v.rule, s)


def _split_phase(v: Variant, s: Stratum, recursive: bool) -> list:
  """[clear temp NEW (recursive)], pipeline A, create-flat-view, pipeline
  B; empty when the rule is compiled unsplit."""
  above = _split_above(v) if v.split_at >= 0 and v.temp_rel else []
  if not above:
    return []
  name = v.rule.name or ""
  if v.delta >= 0:
    name = f"{name}_D{v.delta}"
  a = M.ExecutePipeline(name + "_splitA", above)
  a.clause_order = list(v.clause_order)
  b = M.ExecutePipeline(name + "_splitB", _split_below(v, s, M.NEW if recursive else M.FULL))
  b.clause_order = list(v.clause_order)
  idx = list(range(len(v.temp_vars)))
  ops = [M.ClearRelation(v.temp_rel, M.NEW)] if recursive else []
  return ops + [a, M.CreateFlatView(v.temp_rel, idx, M.NEW), b]


# ------------------------------------------------------ index selection ----
def _complete(idx, arity: int) -> list[int]:
  out = list(idx)
  out += [c for c in range(arity) if c not in out]
  return out


def _indices_of(variants: list[Variant], ver: str | None) -> dict[str, set[tuple]]:
  out: dict[str, set[tuple]] = {}
  for v in variants:
    for a in v.atoms + v.negs:
      if ver is None or a.ver == ver:
        out.setdefault(a.rel, set()).add(tuple(a.index))
  return out


def select_indices(strata: list[Stratum], arity: dict[str, int]) -> None:
  global_map: dict[str, list[list[int]]] = {}
  for s in strata:
    allv = _indices_of(s.base_variants + s.rec_variants, None)
    for rel in sorted(allv):
      lst = global_map.setdefault(rel, [])
      for t in sorted(allv[rel]):
        if list(t) not in lst:
          lst.append(list(t))
  for s in strata:
    if s.recursive:
      dl = _indices_of(s.rec_variants, M.DELTA)
      fl = _indices_of(s.rec_variants, M.FULL)
      groups = [dl, fl]
    else:
      fl = {}
      groups = [_indices_of(s.base_variants, None)]
    for rel in sorted(s.members):
      ar = arity[rel]
      idxs: list[list[int]] = []
      for g in groups:
        for t in sorted(g.get(rel, ())):
          c = _complete(t, ar)
          if c not in idxs:
            idxs.append(c)
      if not idxs:
        for t in global_map.get(rel, []):
          c = _complete(t, ar)
          if c not in idxs:
            idxs.append(c)
      if not idxs:
        idxs.append(list(range(ar)))
      s.required[rel] = idxs
      canon = None
      if s.recursive and rel in fl:
        canon = next((i for i in idxs if len(i) == ar and tuple(i) in fl[rel]), None)
      if canon is None:
        canon = next((i for i in idxs if len(i) == ar), list(range(ar)))
      s.canonical[rel] = list(canon)


# ------------------------------------------------------------- lowering ----
def _lower_body(v: Variant) -> list:
  ops: list = []
  if len(v.atoms) == 1:
    a = v.atoms[0]
    ops.append(M.Scan(list(a.order), a.rel, list(a.index), a.ver, list(a.order[:a.prefix_len])))
  elif len(v.atoms) > 1:
    bound: list[str] = []
    chosen: dict[int, list[int]] = {}
    for var in v.var_order:
      if var not in v.join_vars:
        continue
      srcs = []
      for a in v.atoms:
        if var in a.order:
          chosen[a.clause] = list(a.index)
          srcs.append(M.ColumnSource(a.rel, list(a.index), a.ver,
                                     [x for x in a.order if x in bound], clause_idx=a.clause))
      if srcs:
        ops.append(M.ColumnJoin(var, srcs))
        bound.append(var)
    pos_vars = set().union(*[set(a.order) for a in v.atoms])
    neg_only = {x for a in v.negs for x in a.order if x not in pos_vars}
    free = [x for x in v.var_order if x not in v.join_vars and x not in neg_only]
    for c in v.rule.body:
      if isinstance(c, Atom):
        for x in c.args:
          if x not in v.join_vars and x not in free and x not in neg_only:
            free.append(x)
    if free:
      fset = set(free)
      srcs, vfs = [], []
      for a in v.atoms:
        if not any(x in fset for x in a.order):
          continue
        if a.clause in chosen:
          idx = list(chosen[a.clause])
        else:
          idx = []
          for b in bound:
            if b in a.order:
              col = a.index[a.order.index(b)]
              if col not in idx:
                idx.append(col)
          idx += [c for c in a.index if c not in idx]
        col_var = {a.index[i]: x for i, x in enumerate(a.order) if i < len(a.index)}
        prefix = []
        for col in idx:
          x = col_var.get(col)
          if x is not None and x in bound and x not in prefix:
            prefix.append(x)
        mine = []
        for col in idx:
          x = col_var.get(col)
          if x is not None and x in fset and x not in prefix and x not in mine:
            mine.append(x)
        srcs.append(M.ColumnSource(a.rel, idx, a.ver, prefix, clause_idx=a.clause))
        vfs.append(mine)
      if srcs:
        ops.append(M.CartesianJoin(free, vfs, srcs))
  for a in v.negs:
    ops.append(M.Negation(a.rel, M.FULL, list(a.index), list(a.order[:a.prefix_len])))
  for c in v.rule.body:
    if isinstance(c, Filter):
      ops.append(M.Filter(list(c.vars), c.code))
  return ops


def _inserts(rule: Rule, s: Stratum) -> list:
  out = []
  for h in rule.heads:
    canon = s.canonical.get(h.rel, list(range(len(h.args))))
    out.append(M.InsertInto(h.rel, M.NEW, list(canon), list(h.args)))
  return out


def _pipeline(v: Variant, s: Stratum) -> M.ExecutePipeline:
  name = v.rule.name or ""
  if v.delta >= 0:
    name = f"{name}_D{v.delta}"
  ep = M.ExecutePipeline(name, _lower_body(v) + _inserts(v.rule, s), count=v.rule.count)
  ep.clause_order = list(v.clause_order)
  return ep


def _simple_maintenance(rel: str, idxs, canon) -> list:
  ops = [M.RebuildIndex(rel, list(canon), M.NEW), M.CheckSize(rel, M.NEW),
         M.ComputeDeltaIndex(rel, list(canon)), M.ClearRelation(rel, M.NEW)]
  for idx in idxs:
    if list(idx) != list(canon):
      ops.append(M.RebuildIndexFromIndex(rel, list(canon), list(idx), M.DELTA))
    ops.append(M.MergeIndex(rel, list(idx)))
  return ops


def _loop_maintenance(rel: str, idxs, canon, full_needed: set[tuple]) -> list:
  ops = [M.RebuildIndex(rel, list(canon), M.NEW), M.ClearRelation(rel, M.DELTA),
         M.CheckSize(rel, M.NEW), M.ComputeDeltaIndex(rel, list(canon)),
         M.ClearRelation(rel, M.NEW)]
  for idx in idxs:
    if list(idx) != list(canon):
      ops.append(M.RebuildIndexFromIndex(rel, list(canon), list(idx), M.DELTA))
    if tuple(idx) == tuple(canon) or tuple(idx) in full_needed:
      ops.append(M.MergeIndex(rel, list(idx)))
  return ops


def lower(strata: list[Stratum], arity: dict[str, int]) -> list[M.Step]:
  steps: list[M.Step] = []
  for s in strata:
    if s.recursive:
      pipes, split_ops = [], []
      for v in s.rec_variants:
        ph = _split_phase(v, s, True)
        if ph:
          split_ops += ph
        else:
          pipes.append(_pipeline(v, s))
      hooks = [M.InjectCppHook(v.rule.name or "", v.rule.debug_code)
               for v in s.rec_variants if v.rule.debug_code]
      ops: list = []
      # unsplit rules first (in parallel), then the split phases in order
      if len(pipes) > 1:
        ops.append(M.ParallelGroup(pipes))
      elif pipes:
        ops.append(pipes[0])
      ops += split_ops
      ops += hooks
      full_reads: dict[str, set[tuple]] = {}
      for v in s.rec_variants:
        for a in v.atoms:
          if a.ver == M.FULL:
            full_reads.setdefault(a.rel, set()).add(tuple(_complete(a.index, arity[a.rel])))
      for rel in sorted(s.members):
        if rel in s.required:
          ops += _loop_maintenance(rel, s.required[rel], s.canonical[rel],
                                   full_reads.get(rel, set()))
      if ops:
        steps.append(M.Step(True, M.FixpointPlan(ops)))
        for rel in sorted(s.members):
          if rel in s.canonical:
            steps.append(M.Step(False, M.PostStratumReconstruct(rel, list(s.canonical[rel]))))
    else:
      pipes, maint, modified, split_ops = [], [], [], []
      for v in s.base_variants:
        ph = _split_phase(v, s, False)
        if ph:
          split_ops += ph
        else:
          pipes.append(_pipeline(v, s))
        for h in v.rule.heads:
          if h.rel not in modified:
            modified.append(h.rel)
          if h.rel in s.required:
            maint += _simple_maintenance(h.rel, s.required[h.rel], s.canonical[h.rel])
      ops = list(split_ops)   # the split phases run first, then the other rules
      if len(pipes) > 1:
        ops.append(M.ParallelGroup(pipes))
      elif pipes:
        ops.append(pipes[0])
      ops += maint
      if ops:
        steps.append(M.Step(False, M.FixpointPlan(ops)))
        for rel in modified:
          if rel in s.canonical:
            steps.append(M.Step(False, M.PostStratumReconstruct(rel, list(s.canonical[rel]))))
        for v in s.base_variants:
          if v.rule.debug_code:
            steps.append(M.Step(False, M.InjectCppHook(v.rule.name or "", v.rule.debug_code)))
  return steps


# ----------------------------------------------------------- MIR passes ----
def _plan_pipelines(body) -> list[M.ExecutePipeline]:
  out = []
  if isinstance(body, M.FixpointPlan):
    for op in body.ops:
      if isinstance(op, M.ExecutePipeline):
        out.append(op)
      elif isinstance(op, M.ParallelGroup):
        out.extend(o for o in op.ops if isinstance(o, M.ExecutePipeline))
  return out


def _needed_indices(body, rel: str, out: set) -> None:
  for ep in _plan_pipelines(body):
    for op in ep.ops:
      if isinstance(op, (M.ColumnJoin, M.CartesianJoin)):
        for src in op.sources:
          if src.rel == rel and src.ver in (M.FULL, M.DELTA):
            out.add(tuple(src.index))


def _rebuilds_before_reconstruct(steps: list[M.Step]) -> list[M.Step]:
  out = []
  for i, st in enumerate(steps):
    out.append(st)
    if not isinstance(st.body, M.PostStratumReconstruct):
      continue
    rel = st.body.rel
    merged: set[tuple] = set()
    for j in range(i - 1, -1, -1):
      b = steps[j].body
      if isinstance(b, M.FixpointPlan) and any(
          d.rel == rel for ep in _plan_pipelines(b) for d in ep.dests()):
        merged = {tuple(o.index) for o in b.ops if isinstance(o, M.MergeIndex) and o.rel == rel}
        break
    needed: set[tuple] = set()
    for j in range(i + 1, len(steps)):
      _needed_indices(steps[j].body, rel, needed)
    for idx in sorted(needed - merged):
      out.append(M.Step(False, M.RebuildIndex(rel, list(idx), M.FULL)))
  return out


def _reorder_sources(steps: list[M.Step]) -> None:
  for st in steps:
    for ep in _plan_pipelines(st.body):
      order = getattr(ep, "clause_order", [])

      def pos(src):
        return order.index(src.clause_idx) if src.clause_idx in order else len(order)

      for op in ep.ops:
        if isinstance(op, M.ColumnJoin) and order:
          op.sources.sort(key=pos)
        elif isinstance(op, M.CartesianJoin) and order:
          pairs = sorted(zip(op.sources, op.var_from_source), key=lambda p: pos(p[0]))
          op.sources = [p[0] for p in pairs]
          op.var_from_source = [p[1] for p in pairs]
      for op in ep.ops:
        if isinstance(op, (M.ColumnJoin, M.CartesianJoin)) and len(op.sources) > 1:
          if op.sources[0].prefix or not any(s.prefix for s in op.sources[1:]):
            continue
          if isinstance(op, M.ColumnJoin):
            op.sources.sort(key=lambda s: 0 if s.prefix else 1)
          else:
            pairs = sorted(zip(op.sources, op.var_from_source), key=lambda p: 0 if p[0].prefix else 1)
            op.sources = [p[0] for p in pairs]
            op.var_from_source = [p[1] for p in pairs]


# ---------------------------------------------------------------- entry ----
def relation_decls(program: Program) -> list[RelDecl]:
  decls = [RelDecl(r.name, r.arity, r.input_file, r.print_size) for r in program.relations]
  # split rules' temp relations, declared after the program's own
  # (reference: hir/split.py TempRelSynthesisPass)
  return decls + [RelDecl(n, a) for n, a in split_temps(program.rules)]


def bind_constants(rules: list[Rule]) -> list[Rule]:
  """A constant argument of a positive body atom becomes a fresh variable
  `K<n>_<value>` with an equality filter right after the atom, so the join
  binds the column like any other and the filter keeps the matching rows;
  a user plan's var_order gets the fresh variable after the atom's last
  ordered variable.  Constants in negated atoms (a prefix of the negation's
  index) and in heads (emitted as literals) stay as they are."""
  import dataclasses
  counter = 0
  out = []
  for r in rules:
    body: list = []
    fresh_after: dict[str, list[str]] = {}   # user variable -> fresh ones to place after it
    fresh_lead: list[str] = []
    for c in r.body:
      if not isinstance(c, Atom) or not any(is_const(v) for v in c.args):
        body.append(c)
        continue
      args, filters = [], []
      for v in c.args:
        if is_const(v):
          counter += 1
          fv = f"K{counter}_{const_value(v)}"
          args.append(fv)
          filters.append(Filter((fv,), f"return {fv} == {const_value(v)}u;"))
        else:
          args.append(v)
      body.append(dataclasses.replace(c, args=tuple(args)))
      body += filters
      users = [v for v in c.args if not is_const(v)]
      fresh = [f.vars[0] for f in filters]
      if users:
        fresh_after.setdefault(users[-1], []).extend(fresh)
      else:
        fresh_lead.extend(fresh)
    plans = []
    for p in r.plans:
      if p.var_order:
        vo: list[str] = list(fresh_lead)
        for v in p.var_order:
          vo.append(v)
          vo += fresh_after.get(v, [])
        placed = set(vo)
        for fs in fresh_after.values():   # atoms whose variables the plan leaves out
          vo += [f for f in fs if f not in placed]
        p = dataclasses.replace(p, var_order=tuple(vo))
      plans.append(p)
    out.append(dataclasses.replace(r, body=tuple(body), plans=tuple(plans)))
  return out


def rename_wildcards(rules: list[Rule]) -> list[Rule]:
  """Give every `_`-prefixed variable occurrence a fresh name `_gen<N>`, N
  counting from 1 over the program's rules in order, heads before body
  (src/srdatalog/rule_rewrite.py:34-67): two wildcards are never the same
  variable."""
  import dataclasses
  counter = 0

  def atom(a: Atom) -> Atom:
    nonlocal counter
    args = []
    for v in a.args:
      if v.startswith("_"):
        counter += 1
        v = f"_gen{counter}"
      args.append(v)
    return dataclasses.replace(a, args=tuple(args))

  out = []
  for r in rules:
    heads = tuple(atom(h) for h in r.heads)
    body = []
    for c in r.body:
      if isinstance(c, Atom):
        body.append(atom(c))
      elif isinstance(c, Negation):
        body.append(dataclasses.replace(c, atom=atom(c.atom)))
      else:
        body.append(c)
    out.append(dataclasses.replace(r, heads=heads, body=tuple(body)))
  return out


def compile_to_mir(program: Program) -> M.Program:
  arity = {r.name: r.arity for r in program.relations}
  for n, a in split_temps(program.rules):
    arity[n] = a
  for r in program.rules:
    for c in r.body:
      if not isinstance(c, (Atom, Negation, Filter, Split)):
        raise CompileError(f"rule {r.name}: unsupported body clause {type(c).__name__}")
    if sum(1 for c in r.body if isinstance(c, Split)) > 1:
      raise CompileError(f"rule {r.name}: more than one split marker")
  strata = stratify(rename_wildcards(bind_constants(list(program.rules))))
  for s in strata:
    if s.recursive:
      for r in s.rules:
        for i, c in enumerate(r.body):
          if isinstance(c, Atom) and c.rel in s.members:
            vers = [M.FULL] * len(r.body)
            vers[i] = M.DELTA
            s.rec_variants.append(Variant(r, i, vers))
    else:
      for r in s.rules:
        s.base_variants.append(Variant(r, -1, [M.FULL] * len(r.body)))
    for v in s.base_variants + s.rec_variants:
      plan_variant(v)
  select_indices(strata, arity)
  steps = lower(strata, arity)
  steps = _rebuilds_before_reconstruct(steps)
  _reorder_sources(steps)
  return M.Program(steps)
