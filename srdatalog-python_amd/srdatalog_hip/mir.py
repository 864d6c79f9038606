"""MIR: the imperative plan the MI355X backend executes.

The node set and the S-expression text form are the reference's contract
(src/srdatalog/mir/types.py node kinds; the text form is what its
tests/fixtures/**/*.mir.sexpr goldens hold).  ``print_program`` reproduces
that text byte-for-byte and ``parse_program`` reads it back, so a plan
produced by the reference pipeline can be fed to this backend unchanged.
"""

from __future__ import annotations

from dataclasses import dataclass, field

FULL, DELTA, NEW = "FULL", "DELTA", "NEW"


# --------------------------------------------------------------- nodes ----
@dataclass
class ColumnSource:
  rel: str
  index: list[int]
  ver: str
  prefix: list[str]
  clause_idx: int = field(default=-1, compare=False)   # body clause it came from (not printed)


@dataclass
class Scan:
  vars: list[str]
  rel: str
  index: list[int]
  ver: str
  prefix: list[str] = field(default_factory=list)


@dataclass
class ColumnJoin:
  var: str
  sources: list[ColumnSource]


@dataclass
class CartesianJoin:
  vars: list[str]
  var_from_source: list[list[str]]
  sources: list[ColumnSource]


@dataclass
class Filter:
  vars: list[str]
  code: str


@dataclass
class Negation:
  rel: str
  ver: str
  index: list[int]
  prefix: list[str]


@dataclass
class InsertInto:
  rel: str
  ver: str
  index: list[int]       # dedup index
  terms: list[str]


@dataclass
class ExecutePipeline:
  rule: str
  ops: list            # Scan | ColumnJoin | CartesianJoin | Filter | Negation | InsertInto
  count: bool = False  # .with_count(): count only, no materialisation

  def sources(self) -> list[ColumnSource | Scan | Negation]:
    out = []
    for op in self.ops:
      if isinstance(op, (ColumnJoin, CartesianJoin)):
        out.extend(op.sources)
      elif isinstance(op, (Scan, Negation)):
        out.append(op)
    return out

  def dests(self) -> list[InsertInto]:
    return [op for op in self.ops if isinstance(op, InsertInto)]


@dataclass
class ParallelGroup:
  ops: list[ExecutePipeline]


@dataclass
class RebuildIndex:
  rel: str
  index: list[int]
  ver: str


@dataclass
class ClearRelation:
  rel: str
  ver: str


@dataclass
class CheckSize:
  rel: str
  ver: str


@dataclass
class ComputeDeltaIndex:
  rel: str
  canonical: list[int]


@dataclass
class MergeIndex:
  rel: str
  index: list[int]


@dataclass
class RebuildIndexFromIndex:
  rel: str
  source: list[int]
  target: list[int]
  ver: str


@dataclass
class CreateFlatView:
  """Expose the rows a split rule's first pipeline wrote into its temp
  relation as an index the second pipeline scans
  (src/srdatalog/mir/types.py CreateFlatView)."""
  rel: str
  index: list[int]
  ver: str


@dataclass
class InjectCppHook:
  rule: str
  code: str


@dataclass
class PostStratumReconstruct:
  rel: str
  canonical: list[int]


@dataclass
class FixpointPlan:
  ops: list


@dataclass
class Step:
  recursive: bool
  body: FixpointPlan | PostStratumReconstruct | InjectCppHook


@dataclass
class Program:
  steps: list[Step]

  def pipelines(self) -> list[ExecutePipeline]:
    out = []
    for st in self.steps:
      if isinstance(st.body, FixpointPlan):
        for op in st.body.ops:
          if isinstance(op, ExecutePipeline):
            out.append(op)
          elif isinstance(op, ParallelGroup):
            out.extend(op.ops)
    return out


# ------------------------------------------------------------- printer ----
def _vars(vs) -> str:
  return "(" + " ".join(vs) + ")"


def _idx(rel, cols) -> str:
  return "(" + rel + " " + " ".join(str(c) for c in cols) + ")"


def _cols(cols) -> str:
  return "(" + " ".join(str(c) for c in cols) + ")"


def _spec(rel, cols, ver) -> str:
  return f"(index-spec :schema {rel} :index {_cols(cols)} :ver {ver})"


def _node(n, ind: int) -> str:
  p = " " * ind
  if isinstance(n, ColumnSource):
    return f"{p}(column-source :index {_idx(n.rel, n.index)} :ver {n.ver} :prefix {_vars(n.prefix)})"
  if isinstance(n, Scan):
    return (f"{p}(scan :vars {_vars(n.vars)} :index {_idx(n.rel, n.index)} :ver {n.ver}"
            f" :prefix {_vars(n.prefix)})")
  if isinstance(n, ColumnJoin):
    srcs = "".join(_node(s, ind + 4) + "\n" for s in n.sources)
    return f"{p}(column-join :var {n.var}\n{p}  :sources (\n{srcs}{p}  ))"
  if isinstance(n, CartesianJoin):
    head = f"{p}(cartesian-join :vars {_vars(n.vars)}"
    if n.var_from_source:
      head += " :var-from-source (" + " ".join(_vars(v) for v in n.var_from_source) + ")"
    srcs = "".join(_node(s, ind + 4) + "\n" for s in n.sources)
    return f"{head}\n{p}  :sources (\n{srcs}{p}  ))"
  if isinstance(n, Filter):
    return f'{p}(filter :vars {_vars(n.vars)} :code "{n.code}")'
  if isinstance(n, Negation):
    return (f"{p}(negation :schema {n.rel} :ver {n.ver} :index {_idx(n.rel, n.index)}"
            f" :prefix {_vars(n.prefix)})")
  if isinstance(n, InsertInto):
    return (f"{p}(insert-into :schema {n.rel} :ver {n.ver} :dedup-index {_cols(n.index)}"
            f" :terms {_vars(n.terms)})")
  if isinstance(n, ExecutePipeline):
    src_specs = [_spec(s.rel, s.index, s.ver) for s in n.sources()]
    dst_specs = [_spec(d.rel, d.index, FULL) for d in n.dests()]
    q = " " * (ind + 4)
    out = f"{p}(execute-pipeline :rule {n.rule}\n"
    out += f"{p}  :sources " + ("(tuple\n" + q + ("\n" + q).join(src_specs) + ")" if src_specs
                                else "(tuple)") + "\n"
    out += f"{p}  :dests " + ("(tuple\n" + q + ("\n" + q).join(dst_specs) + ")" if dst_specs
                              else "(tuple)") + "\n"
    for op in n.ops:
      out += _node(op, ind + 4) + "\n"
    return out + f"{p})"
  if isinstance(n, ParallelGroup):
    out = f"{p}(parallel-group  ;; {len(n.ops)} independent ops\n"
    for op in n.ops:
      out += _node(op, ind + 4) + "\n"
    return out + f"{p})"
  if isinstance(n, RebuildIndex):
    return f"{p}(rebuild-index :index {_idx(n.rel, n.index)} :ver {n.ver})"
  if isinstance(n, ClearRelation):
    return f"{p}(clear-relation :schema {n.rel} :ver {n.ver})"
  if isinstance(n, CheckSize):
    return f"{p}(check-size :schema {n.rel} :ver {n.ver})"
  if isinstance(n, ComputeDeltaIndex):
    return f"{p}(compute-delta-index :schema {n.rel} :canonical-index {_cols(n.canonical)})"
  if isinstance(n, MergeIndex):
    return f"{p}(merge-index :index {_idx(n.rel, n.index)})"
  if isinstance(n, RebuildIndexFromIndex):
    return (f"{p}(rebuild-index-from-index :source {_idx(n.rel, n.source)}"
            f" :target {_idx(n.rel, n.target)} :ver {n.ver})")
  if isinstance(n, CreateFlatView):
    return f"{p}(create-flat-view :index {_idx(n.rel, n.index)} :ver {n.ver})"
  if isinstance(n, InjectCppHook):
    return f'{p}(inject-cpp-hook :rule {n.rule} :code "...")'
  if isinstance(n, PostStratumReconstruct):
    return f"{p}(post-stratum-reconstruct-intern-cols :rel {n.rel} :canonical-index {_cols(n.canonical)})"
  if isinstance(n, FixpointPlan):
    out = f"{p}(fixpoint-plan\n"
    for op in n.ops:
      out += _node(op, ind + 4) + "\n"
    return out + f"{p})"
  raise TypeError(f"cannot print MIR node {type(n).__name__}")


def print_program(prog: Program) -> str:
  """The reference's MIR S-expression text (no trailing newline)."""
  out = "(program\n"
  for st in prog.steps:
    out += f"  (step :recursive {'true' if st.recursive else 'false'}\n"
    out += _node(st.body, 8) + "\n"
    out += "  )\n"
  return out + ")"


# -------------------------------------------------------------- parser ----
class _Sym(str):
  """A bare symbol (keywords, names, numbers); plain ``str`` is a string literal."""


_LP, _RP = object(), object()


def _tokenize(text: str):
  i, n = 0, len(text)
  while i < n:
    c = text[i]
    if c.isspace():
      i += 1
    elif c == ";":
      while i < n and text[i] != "\n":
        i += 1
    elif c == "(":
      yield _LP
      i += 1
    elif c == ")":
      yield _RP
      i += 1
    elif c == '"':
      j = i + 1
      buf = []
      while j < n and text[j] != '"':
        if text[j] == "\\" and j + 1 < n:
          j += 1
        buf.append(text[j])
        j += 1
      yield "".join(buf)
      i = j + 1
    else:
      j = i
      while j < n and not text[j].isspace() and text[j] not in "()":
        j += 1
      yield _Sym(text[i:j])
      i = j


def read_sexprs(text: str) -> list:
  """Nested lists of _Sym / str atoms."""
  stack = [[]]
  for tok in _tokenize(text):
    if tok is _LP:
      stack.append([])
    elif tok is _RP:
      if len(stack) == 1:
        raise ValueError("unbalanced S-expression")
      done = stack.pop()
      stack[-1].append(done)
    else:
      stack[-1].append(tok)
  if len(stack) != 1:
    raise ValueError("unbalanced S-expression")
  return stack[0]


def _kw(lst: list, key: str, default=None):
  for i in range(len(lst) - 1):
    if isinstance(lst[i], _Sym) and lst[i] == key:
      return lst[i + 1]
  return default


def _positional(lst: list, start: int = 1) -> list:
  out, i = [], start
  while i < len(lst):
    if isinstance(lst[i], _Sym) and lst[i].startswith(":"):
      i += 2
      continue
    out.append(lst[i])
    i += 1
  return out


def _ints(x) -> list[int]:
  return [int(v) for v in x]


def _strs(x) -> list[str]:
  return [str(v) for v in x]


def _index_ref(x):
  return str(x[0]), _ints(x[1:])


def _parse_node(x):
  head = str(x[0])
  if head == "column-source":
    rel, idx = _index_ref(_kw(x, ":index"))
    return ColumnSource(rel, idx, str(_kw(x, ":ver")), _strs(_kw(x, ":prefix", [])))
  if head == "scan":
    rel, idx = _index_ref(_kw(x, ":index"))
    return Scan(_strs(_kw(x, ":vars")), rel, idx, str(_kw(x, ":ver")), _strs(_kw(x, ":prefix", [])))
  if head == "column-join":
    return ColumnJoin(str(_kw(x, ":var")), [_parse_node(s) for s in _kw(x, ":sources")])
  if head == "cartesian-join":
    vfs = _kw(x, ":var-from-source", [])
    return CartesianJoin(_strs(_kw(x, ":vars")), [_strs(v) for v in vfs],
                         [_parse_node(s) for s in _kw(x, ":sources")])
  if head == "filter":
    return Filter(_strs(_kw(x, ":vars")), str(_kw(x, ":code")))
  if head == "negation":
    _, idx = _index_ref(_kw(x, ":index"))
    return Negation(str(_kw(x, ":schema")), str(_kw(x, ":ver")), idx, _strs(_kw(x, ":prefix", [])))
  if head == "insert-into":
    return InsertInto(str(_kw(x, ":schema")), str(_kw(x, ":ver")), _ints(_kw(x, ":dedup-index")),
                      _strs(_kw(x, ":terms")))
  if head == "execute-pipeline":
    ops = [_parse_node(o) for o in _positional(x) if isinstance(o, list)]
    cnt = _kw(x, ":count")
    return ExecutePipeline(str(_kw(x, ":rule")), ops, count=(cnt == "true"))
  if head == "parallel-group":
    return ParallelGroup([_parse_node(o) for o in _positional(x) if isinstance(o, list)])
  if head == "rebuild-index":
    rel, idx = _index_ref(_kw(x, ":index"))
    return RebuildIndex(rel, idx, str(_kw(x, ":ver")))
  if head == "clear-relation":
    return ClearRelation(str(_kw(x, ":schema")), str(_kw(x, ":ver")))
  if head == "check-size":
    return CheckSize(str(_kw(x, ":schema")), str(_kw(x, ":ver")))
  if head == "compute-delta-index":
    return ComputeDeltaIndex(str(_kw(x, ":schema")), _ints(_kw(x, ":canonical-index")))
  if head == "merge-index":
    rel, idx = _index_ref(_kw(x, ":index"))
    return MergeIndex(rel, idx)
  if head == "rebuild-index-from-index":
    rel, src = _index_ref(_kw(x, ":source"))
    _, tgt = _index_ref(_kw(x, ":target"))
    return RebuildIndexFromIndex(rel, src, tgt, str(_kw(x, ":ver")))
  if head == "create-flat-view":
    rel, idx = _index_ref(_kw(x, ":index"))
    return CreateFlatView(rel, idx, str(_kw(x, ":ver")))
  if head == "inject-cpp-hook":
    return InjectCppHook(str(_kw(x, ":rule")), str(_kw(x, ":code", "")))
  if head == "post-stratum-reconstruct-intern-cols":
    return PostStratumReconstruct(str(_kw(x, ":rel")), _ints(_kw(x, ":canonical-index")))
  if head == "fixpoint-plan":
    return FixpointPlan([_parse_node(o) for o in _positional(x) if isinstance(o, list)])
  raise ValueError(f"unsupported MIR node {head!r}")


def parse_program(text: str) -> Program:
  """Parse the reference's MIR S-expression text into a Program."""
  top = [t for t in read_sexprs(text) if isinstance(t, list) and t and str(t[0]) == "program"]
  if not top:
    raise ValueError("no (program ...) form")
  steps = []
  for st in _positional(top[0]):
    if not (isinstance(st, list) and str(st[0]) == "step"):
      raise ValueError("expected (step ...)")
    rec = str(_kw(st, ":recursive")) == "true"
    body = [b for b in _positional(st) if isinstance(b, list)]
    if len(body) != 1:
      raise ValueError("a step holds exactly one body form")
    steps.append(Step(rec, _parse_node(body[0])))
  return Program(steps)
