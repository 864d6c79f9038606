"""Datalog DSL with the reference's surface (src/srdatalog/dsl.py).

Same names and argument meanings as the reference so programs written for
it run here unchanged::

    x, y, z = Var("x"), Var("y"), Var("z")
    Edge = Relation("Edge", 2, input_file="Edge.csv")
    Path = Relation("Path", 2, print_size=True)
    prog = Program(rules=[
        (Path(x, y) <= Edge(x, y)).named("TCBase"),
        (Path(x, z) <= Path(x, y) & Edge(y, z)).named("TCRec"),
    ])

Covered: Var, Const (and plain ints) as atom arguments, Relation
(input_file / print_size / index_type), atoms, ``&`` conjunction, ``~atom``
negation, Filter, the SPLIT marker, PlanEntry, Rule.named / with_plan /
with_plans / with_count / with_inject_cpp, Program (relations derived from
the rules in first-occurrence order, as in the reference's
_derive_relations, dsl.py:641).  Let and aggregates are not part of the hot
path and are rejected by the compiler.

A constant argument is kept in the atom as the name ``#<value>`` (no
variable name starts with ``#``); the compiler binds it.
"""

from __future__ import annotations

import dataclasses
from dataclasses import dataclass, field


CONST_MARK = "#"


def is_const(name: str) -> bool:
  """Whether an atom argument name is a constant (``#<value>``)."""
  return name.startswith(CONST_MARK)


def const_value(name: str) -> int:
  return int(name[len(CONST_MARK):])


class Const:
  """A constant atom argument (the reference's dsl.Const: a value baked in
  at program construction, e.g. a dataset id read from a meta file).  Values
  are u32 column values; cpp_expr is accepted for the same signature and
  must spell the same integer when given."""

  __slots__ = ("value", "cpp_expr")

  def __init__(self, value, cpp_expr: str | None = None):
    if isinstance(value, bool) or not isinstance(value, int):
      raise TypeError(f"Const: integer values only, got {value!r}")
    if not 0 <= value < (1 << 32):
      raise ValueError(f"Const: {value} is not a u32 column value")
    if cpp_expr is not None and cpp_expr.rstrip("uU") != str(value):
      raise ValueError(f"Const: cpp_expr {cpp_expr!r} does not spell {value}")
    self.value = value
    self.cpp_expr = cpp_expr if cpp_expr is not None else str(value)

  def __repr__(self) -> str:
    return f"Const({self.value})"


class Var:
  __slots__ = ("name",)

  def __init__(self, name: str):
    self.name = name

  def __repr__(self) -> str:
    return f"Var({self.name!r})"


@dataclass(frozen=True)
class Atom:
  rel: str
  args: tuple[str, ...]                 # variable names
  relation: "Relation | None" = field(default=None, compare=False, repr=False)

  def __and__(self, other) -> "Conjunction":
    return Conjunction((self,)) & other

  def __invert__(self) -> "Negation":
    return Negation(self)

  def __le__(self, body) -> "Rule":
    if isinstance(body, Conjunction):
      clauses = body.clauses
    elif isinstance(body, (Atom, Negation, Filter)):
      clauses = (body,)
    else:
      raise TypeError(f"rule body must be atoms / filters, got {type(body).__name__}")
    return Rule(heads=(self,), body=tuple(clauses))


@dataclass(frozen=True)
class Negation:
  atom: Atom

  def __and__(self, other) -> "Conjunction":
    return Conjunction((self,)) & other


@dataclass(frozen=True)
class Filter:
  """Inline C++ predicate over bound variables, e.g.
  ``Filter(("x", "y"), "return x != y;")``; variables are uint32_t."""

  vars: tuple[str, ...]
  code: str

  def __init__(self, vars, code: str):
    object.__setattr__(self, "vars", tuple(v.name if isinstance(v, Var) else str(v) for v in vars))
    object.__setattr__(self, "code", code)

  def __and__(self, other) -> "Conjunction":
    return Conjunction((self,)) & other


@dataclass(frozen=True)
class Split:
  """Split marker: the clauses before it run as a first pipeline writing the
  variables the rest needs into a temp relation ``_temp_<rule name>``; a
  second pipeline scans the temp and joins the clauses after it (negation
  pushdown).  At most one per rule body (reference: dsl.py Split / SPLIT)."""

  def __and__(self, other) -> "Conjunction":
    return Conjunction((self,)) & other


SPLIT = Split()


@dataclass(frozen=True)
class Conjunction:
  clauses: tuple

  def __and__(self, other) -> "Conjunction":
    if isinstance(other, Conjunction):
      return Conjunction(self.clauses + other.clauses)
    if isinstance(other, (Atom, Negation, Filter, Split)):
      return Conjunction(self.clauses + (other,))
    raise TypeError(f"cannot conjoin {type(other).__name__}")


@dataclass(frozen=True)
class PlanEntry:
  """User plan for one variant (delta = body index of the DELTA atom, -1 = base)."""

  delta: int = -1
  var_order: tuple[str, ...] = ()
  clause_order: tuple[int, ...] = ()
  fanout: bool = False
  work_stealing: bool = False
  block_group: bool = False
  dedup_hash: bool = False
  balanced_root: tuple[str, ...] = ()
  balanced_sources: tuple[str, ...] = ()


@dataclass(frozen=True)
class Rule:
  heads: tuple[Atom, ...]
  body: tuple
  name: str | None = None
  plans: tuple[PlanEntry, ...] = ()
  count: bool = False
  debug_code: str = ""

  @property
  def head(self) -> Atom:
    return self.heads[0]

  def named(self, name: str) -> "Rule":
    return dataclasses.replace(self, name=name)

  def with_plan(self, *, delta: int = -1, var_order=None, clause_order=None, fanout=False,
                work_stealing=False, block_group=False, dedup_hash=False, balanced_root=None,
                balanced_sources=None) -> "Rule":
    entry = PlanEntry(delta=delta,
                      var_order=tuple(var_order) if var_order is not None else (),
                      clause_order=tuple(clause_order) if clause_order is not None else (),
                      fanout=fanout, work_stealing=work_stealing, block_group=block_group,
                      dedup_hash=dedup_hash,
                      balanced_root=tuple(balanced_root or ()),
                      balanced_sources=tuple(balanced_sources or ()))
    return dataclasses.replace(self, plans=self.plans + (entry,))

  def with_plans(self, entries) -> "Rule":
    return dataclasses.replace(self, plans=tuple(entries))

  def with_count(self) -> "Rule":
    return dataclasses.replace(self, count=True)

  def with_inject_cpp(self, code: str) -> "Rule":
    return dataclasses.replace(self, debug_code=code)


class Relation:
  """A relation declaration; calling it builds an atom."""

  __slots__ = ("name", "arity", "column_types", "input_file", "print_size", "output_file",
               "index_type", "semiring")

  def __init__(self, name: str, arity: int, column_types=None, *, input_file: str = "",
               print_size: bool = False, output_file: str = "", index_type: str = "",
               semiring: str = "NoProvenance"):
    self.name = name
    self.arity = arity
    self.column_types = column_types or tuple([int] * arity)
    self.input_file = input_file
    self.print_size = print_size
    self.output_file = output_file
    self.index_type = index_type
    self.semiring = semiring

  def __call__(self, *args) -> Atom:
    if len(args) != self.arity:
      raise ValueError(f"{self.name} expects arity {self.arity}, got {len(args)}")
    names = []
    for a in args:
      if isinstance(a, Var):
        names.append(a.name)
      elif isinstance(a, Const) or (isinstance(a, int) and not isinstance(a, bool)):
        c = a if isinstance(a, Const) else Const(a)
        names.append(f"{CONST_MARK}{c.value}")
      else:
        raise TypeError(f"{self.name}: arguments are variables or integer constants, got {a!r}")
    return Atom(self.name, tuple(names), relation=self)

  def __repr__(self) -> str:
    return f"Relation({self.name!r}, arity={self.arity})"


@dataclass
class Program:
  rules: list[Rule] = field(default_factory=list)
  relations: list[Relation] = field(init=False)

  def __post_init__(self) -> None:
    self.relations = _derive_relations(self.rules)

  def add(self, *items: Rule) -> "Program":
    self.rules.extend(items)
    self.relations = _derive_relations(self.rules)
    return self


def _derive_relations(rules) -> list[Relation]:
  out, seen = [], set()

  def take(rel):
    if rel is not None and rel.name not in seen:
      seen.add(rel.name)
      out.append(rel)

  for r in rules:
    for h in r.heads:
      take(h.relation)
    for c in r.body:
      if isinstance(c, Atom):
        take(c.relation)
      elif isinstance(c, Negation):
        take(c.atom.relation)
  return out
