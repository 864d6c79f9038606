"""MIR -> HIP (gfx950) emitter.

For every ``execute-pipeline`` of a MIR program this emits one root kernel
and one slice kernel (instantiated for count and write), plus the host
launchers and the ``SrdPipelineDesc`` table that the runtime's MIR executor
drives (csrc/srd_pipeline.h).  The program's MIR text is embedded verbatim
and the six ``srdatalog_*`` entry points of the reference's generated shim
(src/srdatalog/codegen/jit/main_file.py:656-781) are emitted on top of the
runtime engine.

Kernel shape (wave64, see csrc/srd_join.h):

* level 0  -- the root column-join (or a scan): one thread per distinct key
  of the smallest root source finds the key in every root source and sizes
  level 1 (``key_work``);
* level 1  -- a column-join (candidates from the smallest source range, the
  others binary-searched) or a cartesian product (mixed-radix decode),
  flattened over all keys: one item per lane;
* level 2  -- optional second column-join or cartesian product: each lane's
  binding is published in LDS and the wave walks the concatenated level-2
  spaces 64 items at a time.

This mirrors the join semantics of the reference's emitted kernels
(src/srdatalog/codegen/jit/pipeline.py:360, root.py, instructions.py):
leapfrog intersection of sorted column ranges, prefix lookups, cartesian
expansion, filters and insert-into, with a wave64-native work layout.
"""

from __future__ import annotations

import os

import re
from dataclasses import dataclass, field

from . import mir as M
from .dsl import const_value, is_const

VER_CODE = {M.FULL: "SRD_FULL", M.DELTA: "SRD_DELTA", M.NEW: "SRD_NEW"}


class CodegenError(NotImplementedError):
  pass


def cident(s: str) -> str:
  return re.sub(r"[^A-Za-z0-9_]", "_", s)


@dataclass
class Rg:
  """A row range of one view at a column depth (C++ expressions)."""

  view: int
  depth: int
  b: str
  e: str
  full0: bool = False     # the whole index at depth 0 (distinct values = root_vals)
  uniform: bool = False   # wave-uniform (not per lane)
  bm: str = ""            # heavy-key bitmap slot of the range (~0u: none), when eligible


@dataclass
class PipeInfo:
  ep: M.ExecutePipeline
  views: list[tuple[str, tuple[int, ...], str]] = field(default_factory=list)

  def view_id(self, rel: str, index, ver: str) -> int:
    key = (rel, tuple(index), ver)
    if key not in self.views:
      if len(self.views) >= 8:
        raise CodegenError("more than 8 distinct source indexes in one pipeline")
      self.views.append(key)
    return self.views.index(key)


class PipelineGen:
  """Emits the kernels of one pipeline."""

  def __init__(self, ep: M.ExecutePipeline, arity: dict[str, int], uid: int,
               coop_search: bool = False, ring: bool = True, dyn: bool = False,
               cart_tiles: bool = True):
    self.ep = ep
    # large two-source cartesian products at a nested last level are walked
    # in register tiles instead of the flattened per-item decode
    self.cart_tiles = cart_tiles
    # also emit the device-sized launches (pipelines of recursive steps)
    self.dyn = dyn and not ep.count
    # run-narrowed cooperative searches (srdj::seg_lower_bound) instead of one
    # independent binary search per lane; measured slower on the triangle
    # join, whose searches are latency-bound, so off by default
    self.coop_search = coop_search
    self.arity = arity
    self.ns = f"srdp_{uid}_{cident(ep.rule)}"
    self.info = PipeInfo(ep)
    self.ops = list(ep.ops)
    for op in self.ops:
      if isinstance(op, M.Negation) and not op.prefix:
        raise CodegenError(f"{ep.rule}: negation without bound variables")
      if not isinstance(op, (M.Scan, M.ColumnJoin, M.CartesianJoin, M.Filter, M.Negation,
                             M.InsertInto)):
        raise CodegenError(f"{ep.rule}: unsupported op {type(op).__name__}")
    self.root = self.ops[0]
    if not isinstance(self.root, (M.Scan, M.ColumnJoin)):
      raise CodegenError(f"{ep.rule}: pipeline must start with a scan or a column-join")
    self.levels = [op for op in self.ops[1:] if isinstance(op, (M.ColumnJoin, M.CartesianJoin))]
    for i, lv in enumerate(self.levels):
      if isinstance(lv, M.CartesianJoin) and i != len(self.levels) - 1:
        raise CodegenError(f"{ep.rule}: cartesian join must be the last join level")
    # per-lane checks applied as soon as their variables are bound: filters
    # and negations (no row of the negated index has the bound prefix)
    self.filters = [op for op in self.ops if isinstance(op, (M.Filter, M.Negation))]
    self.dests = [op for op in self.ops if isinstance(op, M.InsertInto)]
    if not self.dests:
      raise CodegenError(f"{ep.rule}: pipeline without insert-into")
    if len(self.dests) > 2:
      raise CodegenError(f"{ep.rule}: more than two insert-into")
    # view ids in source order
    for s in ep.sources():
      self.info.view_id(s.rel, s.index, s.ver)
    # fused mode stages its rows through a per-wave LDS ring (decided in
    # gen_run from the LDS the level frames leave)
    self.ring = False
    self.ring_ok = ring

  # ------------------------------------------------------------ helpers --
  def vid(self, s) -> int:
    return self.info.view_id(s.rel, s.index, s.ver)

  def view_arity(self, v: int) -> int:
    rel, index, _ = self.info.views[v]
    return len(index)

  @staticmethod
  def var(name: str) -> str:
    if is_const(name):   # a constant argument: its literal
      return f"{const_value(name)}u"
    return "v_" + cident(name)

  def col_expr(self, v: int, depth: int) -> str:
    return f"srdj::colp(V.v[{v}], {depth})"

  def size_expr(self, r: Rg, rows: bool) -> str:
    """Iteration size: distinct values (column join) or rows (cartesian)."""
    if r.full0:
      return f"V.v[{r.view}].n" if rows else f"V.v[{r.view}].r"
    return f"({r.e} - {r.b})"

  def find_expr(self, r: Rg, val: str) -> str:
    """Rng of rows in r whose column r.depth equals val."""
    if r.full0:
      return f"srd::root_find(V.v[{r.view}], {val})"
    last = "true" if r.depth == self.view_arity(r.view) - 1 else "false"
    return f"srd::child_find(V.v[{r.view}], {r.depth}, srd::Rng{{{r.b}, {r.e}}}, {val}, {last})"

  # Filters are applied as soon as all their vars are bound.
  def filters_ready(self, bound: set[str], done: set[int]) -> list[tuple[int, M.Filter]]:
    out = []
    for i, f in enumerate(self.filters):
      vs = f.prefix if isinstance(f, M.Negation) else f.vars
      if i not in done and all(v in bound or is_const(v) for v in vs):
        out.append((i, f))
    return out

  def filter_code(self, f) -> str:
    if isinstance(f, M.Negation):
      args = ", ".join(self.var(v) for v in f.prefix)
      return f"!srdj::has_prefix(V.v[{self.vid(f)}], {args})"
    params = ", ".join(f"uint32_t {cident(v)}" for v in f.vars)
    args = ", ".join(self.var(v) for v in f.vars)
    return f"([&]({params}) -> bool {{ {f.code} }})({args})"

  def emit_code(self, cond: str, ind: str) -> str:
    """Wave-cooperative emission of this lane's tuple when `cond` holds.

    MODE 0 counts, 1 writes at the slice's scanned offset, 2 appends to the
    slice's staging segment, 3 writes at the item's own row (direct), 5
    counts the searches only (probe runs), 6 sets the tuple's bit in the
    destination's dense bitmap (a plain load first: most derivations of a
    deduplicating join are repeats, whose bit is already set), and, for a
    bit it sets, its block's bit in the summary the extraction reads."""
    lines = [f"{ind}{{",
             f"{ind}  const uint64_t m_ = srd::ballot({cond});",
             f"{ind}  const uint32_t c_ = __popcll(m_);",
             f"{ind}  if (MODE == 1 && ({cond})) {{",
             f"{ind}    const uint64_t o_ = opos + srd::popc_below(m_);"]
    for di, d in enumerate(self.dests):
      for c, relcol in enumerate(d.index):
        lines.append(f"{ind}    A.out[{di}].cols[{c} * A.out[{di}].stride + ob{di}_ + o_]"
                     f" = {self.var(d.terms[relcol])};")
    lines += [f"{ind}  }}",
              f"{ind}  if (MODE == 3 && ({cond})) {{"]
    for di, d in enumerate(self.dests):
      for c, relcol in enumerate(d.index):
        lines.append(f"{ind}    A.out[{di}].cols[{c} * A.out[{di}].stride + ob{di}_ + item]"
                     f" = {self.var(d.terms[relcol])};")
    lines += [f"{ind}  }}",
              f"{ind}  if (MODE == 2 && c_) {{",
              f"{ind}    if (!seg_dead && (seg_start == ~uint64_t(0) || seg_used + c_ > seg_room)) {{"]
    if self.ring:
      lines.append(f"{ind}      if (seg_start != ~uint64_t(0)) ring_out(seg_start + seg_used, false);")
    lines += [f"{ind}      srdj::seg_close(A, slice, seg_start, seg_used, seg_local);",
              f"{ind}      seg_start = srdj::seg_open(A);",
              f"{ind}      seg_room = SRD_SEG;",
              f"{ind}      seg_used = 0;",
              f"{ind}      seg_local = cnt;",
              f"{ind}      seg_dead = seg_start == ~uint64_t(0);",
              f"{ind}    }}",
              ]
    if self.ring:
      lines += [f"{ind}    if (seg_start != ~uint64_t(0)) {{",
                f"{ind}      const uint64_t p0_ = seg_start + seg_used;",
                f"{ind}      if ({cond}) {{",
                f"{ind}        const uint32_t r_ = uint32_t(p0_ + srd::popc_below(m_)) & 127u;"]
      j = 0
      for di, d in enumerate(self.dests):
        for c, relcol in enumerate(d.index):
          lines.append(f"{ind}        ring_s[wv][{j}][r_] = {self.var(d.terms[relcol])};")
          j += 1
      lines += [f"{ind}      }}",
                f"{ind}      if (((p0_ + c_) >> 6) != (p0_ >> 6)) ring_out(p0_, true);",
                f"{ind}    }}"]
    else:
      lines += [f"{ind}    if (seg_start != ~uint64_t(0) && ({cond})) {{",
                f"{ind}      const uint64_t o_ = seg_start + seg_used + srd::popc_below(m_);"]
      for di, d in enumerate(self.dests):
        for c, relcol in enumerate(d.index):
          lines.append(f"{ind}      A.stage[{di * 8 + c}ull * A.stage_cap + o_] = {self.var(d.terms[relcol])};")
      lines += [f"{ind}    }}"]
    lines += [f"{ind}    seg_used += c_;",
              f"{ind}  }}"]
    if self.mark_ok():
      d = self.dests[0]
      k = len(d.index)
      parts = [f"(uint64_t({self.var(d.terms[relcol])}) << (A.mark_shift * {k - 1 - c}u))"
               for c, relcol in enumerate(d.index)]
      lines += [f"{ind}  if (MODE == 6 && ({cond})) {{",
                f"{ind}    const uint64_t key_ = {' | '.join(parts)};",
                f"{ind}    uint32_t* w_ = A.mark_bits + (key_ >> 5);",
                f"{ind}    const uint32_t b_ = 1u << uint32_t(key_ & 31u);",
                f"{ind}    if (!(*w_ & b_)) {{",
                f"{ind}      atomicOr(w_, b_);",
                f"{ind}      uint32_t* s_ = A.mark_sum + (key_ >> 23);",
                f"{ind}      const uint32_t sb_ = 1u << uint32_t((key_ >> 18) & 31u);",
                f"{ind}      if (!(*s_ & sb_)) atomicOr(s_, sb_);",
                f"{ind}    }}",
                f"{ind}  }}"]
    lines += [f"{ind}  opos += c_;",
              f"{ind}  cnt += c_;",
              f"{ind}}}"]
    return "\n".join(lines)

  def mark_ok(self) -> bool:
    """Mode 6 applies: one destination whose terms are all variables (every
    value then comes from a source column, so the engine can bound the key
    space), written by a materialising rule."""
    if len(self.dests) != 1 or self.ep.count:
      return False
    d = self.dests[0]
    return 1 <= len(d.index) <= 4 and all(isinstance(d.terms[c], str) and not is_const(d.terms[c])
                                          for c in d.index)

  def direct_ok(self) -> bool:
    """Every level-1 item yields exactly one tuple: no filters, and either a
    scan root, a root join with no nested level, or a cartesian last level
    directly under the root."""
    if self.filters:
      return False
    if isinstance(self.root, M.Scan):
      return not self.levels
    if not self.levels:
      return True
    return len(self.levels) == 1 and isinstance(self.levels[0], M.CartesianJoin)

  # ------------------------------------------------------------ ranges ---
  def sorted_output(self) -> bool:
    """Whether the rule writes its one destination strictly increasing in
    the destination's index order: the root and every level are column
    joins and the written row is exactly (root var, level-1 var, ...).  Keys
    come in ascending order, each level's candidates are the distinct values
    of a sorted range in ascending order, and emission keeps candidate order,
    so rows come out sorted and without repeats in every mode."""
    if isinstance(self.root, M.Scan) or len(self.dests) != 1:
      return False
    if not all(isinstance(lv, M.ColumnJoin) for lv in self.levels):
      return False
    d = self.dests[0]
    row = [d.terms[relcol] for relcol in d.index]
    want = [self.root.var] + [lv.var for lv in self.levels]
    return all(isinstance(t, str) for t in row) and row == want

  def bm_eligible(self, r: Rg) -> bool:
    """A per-lane depth-1 range of a two-column view: its values are the
    last column, so a membership bitmap can replace the search."""
    return (not r.full0 and not r.uniform and r.depth == 1 and self.view_arity(r.view) == 2)

  def bm_views(self) -> set:
    """Views that the second level of the join searches at depth 1."""
    out = set()
    for lv in self.levels[1:2]:
      for s_ in lv.sources:
        if len(s_.prefix) == 1 and self.view_arity(self.vid(s_)) == 2:
          out.add(self.vid(s_))
    return out

  def root_pos_views(self) -> set:
    """Views whose root values a nested join level searches for a key."""
    out = set()
    for lv in self.levels:
      if isinstance(lv, M.ColumnJoin) and len(lv.sources) > 1:
        for s_ in lv.sources:
          if not s_.prefix:
            out.add(self.vid(s_))
    return out

  def resolve(self, reg: dict, v: int, prefix: tuple, code: list[str], ind: str, tag: str) -> Rg:
    """Range of view v with the given bound prefix, narrowing if needed."""
    key = (v, prefix)
    if key in reg:
      return reg[key]
    if not prefix:
      r = Rg(v, 0, "0u", f"V.v[{v}].n", full0=True, uniform=True)
      reg[key] = r
      return r
    parent = self.resolve(reg, v, prefix[:-1], code, ind, tag)
    self._nr = getattr(self, "_nr", 0) + 1
    nm = f"nr_{tag}_{v}_{len(prefix)}_{self._nr}"
    code.append(f"{ind}const srd::Rng {nm} = {self.find_expr(parent, self.var(prefix[-1]))};")
    if getattr(self, "_in_run", False):
      code.append(f"{ind}if constexpr (MODE == 5) probes_ += 1u;")
    r = Rg(v, len(prefix), f"{nm}.b", f"{nm}.e")
    reg[key] = r
    return r

  # -------------------------------------------------------------- root ---
  def root_slots(self):
    """(view, prefix) ranges computed by the root kernel and stored per key."""
    if isinstance(self.root, M.Scan):
      return []
    v0 = self.root.var
    slots = []
    for s in self.root.sources:
      key = (self.vid(s), (v0,))
      if key not in slots:
        slots.append(key)
    return slots

  def gen_root(self) -> str:
    self._in_run = False
    # DYN: sizes and views come from the device-resident A.dyn (device-sized
    # iterations, whose launches are enqueued or captured without knowing
    # them); the grid is fixed and strides over the keys
    out = ["template <bool DYN>",
           f"__global__ __launch_bounds__(256) void root_k(SrdLaunchArgs A) {{",
           "  const SrdViews& V = DYN ? A.dyn->views : A.views;",
           "  if (DYN && A.dyn->skip) return;",
           "  const uint32_t nroot_ = DYN ? A.dyn->n_root : A.n_root;",
           "  const uint32_t drv_ = DYN ? A.dyn->drv_view : A.drv_view;",
           "  (void)drv_;",
           "  const uint32_t kstep_ = DYN ? gridDim.x * 256u : nroot_;",
           "  for (uint32_t k = blockIdx.x * 256u + threadIdx.x; k < nroot_;"
           " k = (nroot_ - k > kstep_) ? k + kstep_ : nroot_) {"]
    if isinstance(self.root, M.Scan):
      out += [f"  A.key_work[k] = V.v[{self.vid(self.root)}].n;", "  A.key_aux[k] = 0;", "  }", "}"]
      return "\n".join(out)
    v0 = self.root.var
    slots = self.root_slots()
    out += ["  const srd::IndexView& D = V.v[drv_];",
            "  const uint32_t key = D.root_vals[k];",
            f"  const uint32_t {self.var(v0)} = key;",
            "  bool ok = true;"]
    reg = {}
    for si, (v, pre) in enumerate(slots):
      out.append(f"  const srd::Rng rs{si} = (drv_ == {v}u) ? srd::Rng{{D.root_offs[k], "
                 f"D.root_offs[k + 1]}} : srd::root_find(V.v[{v}], key);")
      out.append(f"  ok = ok && rs{si}.e > rs{si}.b;")
      reg[(v, pre)] = Rg(v, 1, f"rs{si}.b", f"rs{si}.e")
    out.append("  uint64_t work = 0;")
    out.append("  uint32_t drv = 0;")
    out.append("  if (ok) {")
    body: list[str] = []
    if not self.levels:
      body.append("    work = 1;")
    else:
      lv = self.levels[0]
      rngs = []
      for s in lv.sources:
        rngs.append(self.resolve(reg, self.vid(s), tuple(s.prefix), body, "    ", "r"))
      if isinstance(lv, M.ColumnJoin):
        body.append(f"    uint64_t best = {self.size_expr(rngs[0], False)};")
        for j in range(1, len(rngs)):
          body.append(f"    {{ const uint64_t z = {self.size_expr(rngs[j], False)}; "
                      f"if (z < best) {{ best = z; drv = {j}; }} }}")
        body.append("    work = best;")
      else:
        body.append("    work = 1;")
        for r in rngs:
          body.append(f"    work *= uint64_t({self.size_expr(r, True)});")
    out += body
    out.append("  }")
    # probe runs: root_find of every slot but the driver's, and the range
    # narrowings of a key that has level-1 work
    n_nr = sum(1 for line in body if "const srd::Rng nr_" in line)
    out.append(f"  if (A.probe_ctr) atomicAdd(A.probe_ctr, {len(slots) - 1}ull + (ok ? {n_nr}ull : 0ull));")
    for si in range(len(slots)):
      out.append(f"  A.key_rng[uint64_t(k) * {2 * len(slots)} + {2 * si}] = rs{si}.b;")
      out.append(f"  A.key_rng[uint64_t(k) * {2 * len(slots)} + {2 * si + 1}] = rs{si}.e;")
    out += ["  A.key_aux[k] = drv;", "  A.key_work[k] = work;", "  }", "}"]
    return "\n".join(out)

  # -------------------------------------------------------- level code ---
  def gen_cj_candidate(self, lv: M.ColumnJoin, rngs: list[Rg], drv_expr: str, item: str,
                       hit: str, ind: str, tag: str, active: str, seg: str) -> tuple[list[str], dict]:
    """Code evaluating one candidate of a column join per lane; returns code
    (to be placed in wave-uniform control flow) and the child ranges
    registered for (view, prefix + var).

    Each active lane takes its item's value from the driving source
    (`drv_expr`, per lane); the other sources are then searched one slot at
    a time, slot t being source t or t+1 depending on the lane's driver.
    With cooperative search on, each source is searched instead with
    srdj::seg_lower_bound, which needs every lane of the wave to take part
    (`seg` groups the lanes of one owner)."""
    var = self.var(lv.var)
    code = [f"{ind}uint32_t {var} = 0;", f"{ind}bool {tag}on = false;"]
    child = {}
    for j, (s, r) in enumerate(zip(lv.sources, rngs)):
      code.append(f"{ind}uint32_t {tag}c{j}b = 0, {tag}c{j}e = 0;")
      child[(self.vid(s), tuple(s.prefix) + (lv.var,))] = Rg(self.vid(s), len(s.prefix) + 1,
                                                             f"{tag}c{j}b", f"{tag}c{j}e")
    lasts = {r.depth == self.view_arity(r.view) - 1 for r in rngs if not r.full0}
    if len(rngs) > 1 and not any(r.full0 for r in rngs) and len(lasts) == 1:
      # every source is a column range of one kind: read the driver's value
      # through per-lane selects instead of a switch (no divergent cases)
      last = lasts.pop()
      i2 = ind + "  "

      def pick(vals):
        e = vals[-1]
        for j in range(len(vals) - 2, -1, -1):
          e = f"({drv_expr} == {j}u ? {vals[j]} : {e})"
        return e
      code.append(f"{ind}if ({active}) {{")
      code.append(f"{i2}const uint32_t* col_ = {pick([self.col_expr(r.view, r.depth) for r in rngs])};")
      code.append(f"{i2}const uint32_t b_ = {pick([f'uint32_t({r.b})' for r in rngs])};")
      code.append(f"{i2}const uint32_t pos_ = b_ + uint32_t({item});")
      code.append(f"{i2}const uint32_t val_ = col_[pos_];")
      if last:
        code.append(f"{i2}{tag}on = true;")
        code.append(f"{i2}const srd::Rng cd_ = srd::Rng{{pos_, pos_ + 1}};")
      else:
        code.append(f"{i2}const uint32_t e_ = {pick([f'uint32_t({r.e})' for r in rngs])};")
        code.append(f"{i2}{tag}on = pos_ == b_ || col_[pos_ - 1] != val_;")
        code.append(f"{i2}const srd::Rng cd_ = srdj::run_at(col_, pos_, e_, val_, false);")
        code.append(f"{i2}if constexpr (MODE == 5) probes_ += 1u;")
      code.append(f"{i2}{var} = val_;")
      for j in range(len(rngs)):
        code.append(f"{i2}if ({drv_expr} == {j}u) {{ {tag}c{j}b = cd_.b; {tag}c{j}e = cd_.e; }}")
      code.append(f"{ind}}}")
    else:
      code.append(f"{ind}if ({active}) {{")
      code.append(f"{ind}  switch ({drv_expr}) {{")
      for j, r in enumerate(rngs):
        code.append(f"{ind}    case {j}: {{")
        i2 = ind + "      "
        if r.full0:
          code.append(f"{i2}const uint32_t pos_ = uint32_t({item});")
          code.append(f"{i2}const uint32_t val_ = V.v[{r.view}].root_vals[pos_];")
          code.append(f"{i2}const srd::Rng cd_ = srd::Rng{{V.v[{r.view}].root_offs[pos_], "
                      f"V.v[{r.view}].root_offs[pos_ + 1]}};")
          code.append(f"{i2}const bool first_ = true;")
        else:
          col = self.col_expr(r.view, r.depth)
          last = r.depth == self.view_arity(r.view) - 1
          code.append(f"{i2}const uint32_t* col_ = {col};")
          code.append(f"{i2}const uint32_t pos_ = {r.b} + uint32_t({item});")
          code.append(f"{i2}const uint32_t val_ = col_[pos_];")
          if last:
            code.append(f"{i2}const bool first_ = true;")
            code.append(f"{i2}const srd::Rng cd_ = srd::Rng{{pos_, pos_ + 1}};")
          else:
            code.append(f"{i2}const bool first_ = pos_ == {r.b} || col_[pos_ - 1] != val_;")
            code.append(f"{i2}const srd::Rng cd_ = srdj::run_at(col_, pos_, {r.e}, val_, false);")
            code.append(f"{i2}if constexpr (MODE == 5) probes_ += 1u;")
        code.append(f"{i2}{tag}on = first_;")
        code.append(f"{i2}{var} = val_;")
        code.append(f"{i2}{tag}c{j}b = cd_.b; {tag}c{j}e = cd_.e;")
        code.append(f"{i2}break;")
        code.append(f"{ind}    }}")
      code.append(f"{ind}  }}")
      code.append(f"{ind}}}")
    if len(rngs) > 1 and not self.coop_search:
      # One search per slot t: source t when the driver lies above t, else
      # source t+1, so a wave whose lanes drive from different sources runs
      # each slot's search once (with per-lane column and range) instead of
      # one divergent search per source.
      def src(m):
        rm = rngs[m]
        if rm.full0:
          return f"V.v[{rm.view}].root_vals", "0u", f"V.v[{rm.view}].r"
        return self.col_expr(rm.view, rm.depth), rm.b, rm.e

      def child_code(m, i3):
        rm = rngs[m]
        if rm.full0:
          return [f"{i3}{tag}c{m}b = V.v[{rm.view}].root_offs[lb_]; "
                  f"{tag}c{m}e = V.v[{rm.view}].root_offs[lb_ + 1];"]
        if rm.depth == self.view_arity(rm.view) - 1:
          return [f"{i3}{tag}c{m}b = lb_; {tag}c{m}e = lb_ + 1;"]
        return [f"{i3}{tag}c{m}b = lb_; {tag}c{m}e = srd::upper_bound_la(col_, lb_ + 1, e_, {var});",
                f"{i3}if constexpr (MODE == 5) probes_ += 1u;"]

      for t in range(len(rngs) - 1):
        (ca, ba, ea), (cb, bb, eb) = src(t), src(t + 1)
        i2 = ind + "  "
        code.append(f"{ind}{{  // search slot {t}: source {t} if the driver is above {t}, else {t + 1}")
        code.append(f"{i2}const bool lo_ = {drv_expr} > {t}u;")
        code.append(f"{i2}const uint32_t* col_ = lo_ ? {ca} : {cb};")
        code.append(f"{i2}const uint32_t b_ = lo_ ? uint32_t({ba}) : uint32_t({bb});")
        code.append(f"{i2}const uint32_t e_ = lo_ ? uint32_t({ea}) : uint32_t({eb});")
        ha, hb_ = rngs[t].bm or "~0u", rngs[t + 1].bm or "~0u"
        va, vb = rngs[t].view, rngs[t + 1].view
        code.append(f"{i2}bool hit_ = false;")
        code.append(f"{i2}uint32_t lb_ = 0u;")
        code.append(f"{i2}if constexpr (MODE == 5) probes_ += {tag}on ? 1u : 0u;")
        search = [f"lb_ = {tag}on ? srd::lower_bound_la(col_, b_, e_, {var}) : 0u;",
                  f"hit_ = {tag}on && lb_ < e_ && col_[lb_] == {var};"]
        if rngs[t].full0 or rngs[t + 1].full0:
          # a key looked up among a view's root values: one load from the
          # view's dense root map when it has one
          pa = f"V.v[{va}].root_pos" if rngs[t].full0 else "nullptr"
          pb = f"V.v[{vb}].root_pos" if rngs[t + 1].full0 else "nullptr"
          ma = f"V.v[{va}].root_max" if rngs[t].full0 else "0u"
          mb = f"V.v[{vb}].root_max" if rngs[t + 1].full0 else "0u"
          search = [f"const uint32_t* rp_ = lo_ ? {pa} : {pb};",
                    f"const uint32_t rm_ = lo_ ? {ma} : {mb};",
                    "if (rp_ != nullptr) {",
                    f"  const uint32_t q_ = ({tag}on && {var} <= rm_) ? rp_[{var}] : ~0u;",
                    "  hit_ = q_ != ~0u;",
                    "  lb_ = hit_ ? q_ : 0u;",
                    "} else {"] + ["  " + x for x in search] + ["}"]
        if rngs[t].bm or rngs[t + 1].bm:
          # a heavy key's range: one bitmap probe instead of the search (the
          # range is the view's last column, so its child range is unused)
          code.append(f"{i2}const uint32_t hs_ = lo_ ? {ha} : {hb_};")
          code.append(f"{i2}if (hs_ != ~0u) {{")
          code.append(f"{i2}  hit_ = {tag}on && srdj::heavy_has(lo_ ? V.v[{va}] : V.v[{vb}], hs_, {var});")
          code.append(f"{i2}}} else {{")
          code += [f"{i2}  {x}" for x in search]
          code.append(f"{i2}}}")
        else:
          code += [f"{i2}{x}" for x in search]
        code.append(f"{i2}if ({tag}on) {{")
        code.append(f"{i2}  {tag}on = hit_;")
        code.append(f"{i2}  if (hit_) {{")
        code.append(f"{i2}    if (lo_) {{")
        code += child_code(t, i2 + "      ")
        code.append(f"{i2}    }} else {{")
        code += child_code(t + 1, i2 + "      ")
        code.append(f"{i2}    }}")
        code.append(f"{i2}  }}")
        code.append(f"{i2}}}")
        code.append(f"{ind}}}")
    elif len(rngs) > 1:
      for m, rm in enumerate(rngs):
        code.append(f"{ind}{{  // search source {m}")
        i2 = ind + "  "
        code.append(f"{i2}const bool on_ = {tag}on && {drv_expr} != {m}u;")
        code.append(f"{i2}if constexpr (MODE == 5) probes_ += on_ ? 1u : 0u;")
        if rm.full0:
          code.append(f"{i2}const uint32_t* col_ = V.v[{rm.view}].root_vals;")
          code.append(f"{i2}const uint32_t b_ = 0u, e_ = V.v[{rm.view}].r;")
        else:
          code.append(f"{i2}const uint32_t* col_ = {self.col_expr(rm.view, rm.depth)};")
          code.append(f"{i2}const uint32_t b_ = {rm.b}, e_ = {rm.e};")
        if self.coop_search:
          code.append(f"{i2}const uint32_t lb_ = srdj::seg_lower_bound(col_, b_, e_, {var}, on_, {seg});")
        else:
          code.append(f"{i2}const uint32_t lb_ = on_ ? srd::lower_bound_la(col_, b_, e_, {var}) : 0u;")
        code.append(f"{i2}if (on_) {{")
        code.append(f"{i2}  const bool hit_ = lb_ < e_ && col_[lb_] == {var};")
        code.append(f"{i2}  {tag}on = hit_;")
        if rm.full0:
          code.append(f"{i2}  if (hit_) {{ {tag}c{m}b = V.v[{rm.view}].root_offs[lb_]; "
                      f"{tag}c{m}e = V.v[{rm.view}].root_offs[lb_ + 1]; }}")
        elif rm.depth == self.view_arity(rm.view) - 1:
          code.append(f"{i2}  if (hit_) {{ {tag}c{m}b = lb_; {tag}c{m}e = lb_ + 1; }}")
        else:
          code.append(f"{i2}  if (hit_) {{ {tag}c{m}b = lb_; "
                      f"{tag}c{m}e = srd::upper_bound_la(col_, lb_ + 1, e_, {var}); }}")
          code.append(f"{i2}  if constexpr (MODE == 5) probes_ += hit_ ? 1u : 0u;")
        code.append(f"{i2}}}")
        code.append(f"{ind}}}")
    code.append(f"{ind}{hit} = {tag}on;")
    return code, child

  def gen_cart(self, lv: M.CartesianJoin, rngs: list[Rg], item: str, guard: str, ind: str,
               tag: str) -> list[str]:
    """Mixed-radix decode of a flattened cartesian item (last source fastest);
    reads happen only under `guard`, where every range is non-empty."""
    if not lv.var_from_source or len(lv.var_from_source) != len(rngs):
      raise CodegenError(f"{self.ep.rule}: cartesian join without :var-from-source")
    code = []
    for vs in lv.var_from_source:
      for vname in vs:
        code.append(f"{ind}uint32_t {self.var(vname)} = 0;")
    code.append(f"{ind}if ({guard}) {{")
    i2 = ind + "  "
    code.append(f"{i2}uint64_t rem_{tag} = {item};")
    for j in reversed(range(len(rngs))):
      z = f"uint64_t({self.size_expr(rngs[j], True)})"
      code.append(f"{i2}const uint32_t off{tag}{j} = uint32_t(rem_{tag} % {z});")
      code.append(f"{i2}rem_{tag} /= {z};")
    for j, (vs, r) in enumerate(zip(lv.var_from_source, rngs)):
      code.append(f"{i2}const uint32_t row{tag}{j} = {r.b} + off{tag}{j};")
      for t, vname in enumerate(vs):
        code.append(f"{i2}{self.var(vname)} = {self.col_expr(r.view, r.depth + t)}[row{tag}{j}];")
    code.append(f"{ind}}}")
    return code

  # ------------------------------------------------------- slice kernel ---
  def gen_run(self) -> str:
    self._in_run = True
    o = ["template <int MODE, bool DYN>",
         # eight waves per SIMD: the persistent grid launches 8 blocks per CU,
         # and with more than ~100 SGPRs a wave the eighth block could only
         # start when another finished (the compiler spills SGPRs to VGPR lanes)
         "__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) "
         "void run_k(SrdLaunchArgs A) {",
         "  const SrdViews& V = DYN ? A.dyn->views : A.views;",
         "  if (DYN && A.dyn->skip) return;",
         "  const uint32_t nroot_ = DYN ? A.dyn->n_root : A.n_root;",
         "  const uint32_t drv_ = DYN ? A.dyn->drv_view : A.drv_view;",
         "  (void)drv_;",
         "  const uint64_t tw_ = DYN ? A.dyn->total_work : A.total_work;",
         "  const uint64_t ns_ = DYN ? A.dyn->n_slices : A.n_slices;"]
    for di in range(len(self.dests)):
      o.append(f"  const uint64_t ob{di}_ = DYN ? A.dyn->out_base[{di}] : A.out[{di}].base;")
    o += ["  const uint32_t lane = srd::lane_id();"]
    # levels walked by the wave from LDS frames: every level below the
    # first (the first is one item per lane), or, under a scan root (whose
    # rows are the items), every level
    scan_root = isinstance(self.root, M.Scan)
    deep = len(self.levels) >= 2 or (scan_root and len(self.levels) >= 1)
    self._lds = []
    if deep or self.ring:
      o += ["  const uint32_t wv = threadIdx.x / 64u;"]
    lds_at = len(o)
    ring_cols = sum(len(d.index) for d in self.dests)
    if self.ring:
      # Fused mode: the row for staging position p waits in ring_s[wv][.][p & 127]
      # until its 64-row block is complete, and the block is then written
      # with one full-width store per column.  Rows scattered straight to
      # staging a few at a time left partly written cache lines, which the
      # L2 evicted and wrote back more than once.  Segments start at
      # multiples of SRD_SEG, so blocks never straddle two segments.  The
      # blocks are written non-temporal: staging is read only by the gather
      # after the launch, and cached it would evict the lines the join's
      # searches and bitmap probes reuse.
      self._lds.append(f"  __shared__ uint32_t ring_s[4][{ring_cols}][128];")
      o += ["  // write out the block holding staging position p: all of it (full) or",
            "  // its rows below p (a segment's last, partly filled block)",
            "  auto ring_out = [&](uint64_t p, bool full) {",
            "    const uint64_t b_ = p & ~uint64_t(63);",
            "    const uint32_t n_ = full ? 64u : uint32_t(p & 63u);",
            "    __builtin_amdgcn_wave_barrier();",
            "    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, \"wavefront\");",
            "    if (lane < n_) {",
            "      const uint32_t r_ = uint32_t(b_ + lane) & 127u;"]
      j = 0
      for di, d in enumerate(self.dests):
        for c in range(len(d.index)):
          o.append(f"      __builtin_nontemporal_store(ring_s[wv][{j}][r_], "
                   f"&A.stage[{di * 8 + c}ull * A.stage_cap + b_ + lane]);")
          j += 1
      o += ["    }",
            "    __builtin_amdgcn_wave_barrier();",
            "  };"]
    slots = self.root_slots()
    ns = len(slots)
    o += ["  // fused mode: the wave's open staging segment, kept across slices (a",
          "  // slice's rows in it are one record; the next slice goes on behind them)",
          "  uint32_t probes_ = 0;  // MODE 5: searches, map lookups and bitmap tests of this lane",
          "  uint64_t seg_start = ~uint64_t(0);",
          "  uint32_t seg_room = 0;",
          "  bool seg_dead = false;  // staging exhausted: count only (host re-runs the write pass)",
          "  for (;;) {",
          "    const uint64_t slice = srdj::next_slice(A, ns_);  // wave-uniform",
          "    if (slice >= ns_) break;",
          "    const uint64_t it0 = A.slice_start ? A.slice_start[slice] : slice * A.grain;",
          "    const uint64_t it1 = A.slice_start ? A.slice_start[slice + 1]"
          " : min(it0 + A.grain, tw_);",
          "    uint64_t opos = MODE == 1 ? A.offsets[slice] : 0;",
          "    uint32_t cnt = 0;",
          "    uint32_t seg_used = 0, seg_local = 0;",
          "    uint32_t kc = ~0u;  // key of the previous batch's last item",
          "    for (uint64_t base = it0; base < it1; base += 64) {",
          "      const uint64_t item = base + lane;",
          "      const bool act = item < it1;",
          "      const uint64_t last_item = min(base + 63, it1 - 1);",
          "      const uint32_t klo = kc == ~0u ? srdj::find_slot(A.work_start, nroot_, base)",
          "                                     : srdj::find_slot_from(A.work_start, kc, nroot_, base);",
          "      const uint32_t khi = srdj::find_slot_from(A.work_start, klo, nroot_, last_item) + 1;",
          "      kc = khi - 1;",
          "      uint32_t k = klo;",
          "      if (act && khi - klo > 1) {",
          "        uint32_t lo = klo, hi = khi;",
          "        while (hi - lo > 1) { const uint32_t mid = (lo + hi) >> 1;"
          " if (A.work_start[mid] <= item) lo = mid; else hi = mid; }",
          "        k = lo;",
          "      }",
          "      const uint64_t li = act ? item - A.work_start[k] : 0;"]
    ind = "      "
    bound: set[str] = set()
    done: set[int] = set()
    reg: dict = {}
    if isinstance(self.root, M.Scan):
      sv = self.vid(self.root)
      o.append(f"{ind}const uint32_t row0 = uint32_t(li);")
      for t, vname in enumerate(self.root.vars):
        o.append(f"{ind}const uint32_t {self.var(vname)} = act ? {self.col_expr(sv, t)}[row0] : 0u;")
        bound.add(vname)
      o.append(f"{ind}bool h1 = act;")
    else:
      v0 = self.root.var
      o.append(f"{ind}const uint32_t {self.var(v0)} = act ? V.v[drv_].root_vals[k] : 0u;")
      bound.add(v0)
      for si, (v, pre) in enumerate(slots):
        o.append(f"{ind}const uint32_t rs{si}b = act ? A.key_rng[uint64_t(k) * {2 * ns} + {2 * si}] : 0u;")
        o.append(f"{ind}const uint32_t rs{si}e = act ? A.key_rng[uint64_t(k) * {2 * ns} + {2 * si + 1}] : 0u;")
        reg[(v, pre)] = Rg(v, 1, f"rs{si}b", f"rs{si}e")
      o.append(f"{ind}const uint32_t l1drv = act ? A.key_aux[k] : 0u;")
      o.append(f"{ind}bool h1 = false;")
      if not self.levels:
        o.append(f"{ind}h1 = act;")
      else:
        lv = self.levels[0]
        rngs = [self.resolve(reg, self.vid(s), tuple(s.prefix), o, ind, "a") for s in lv.sources]
        if isinstance(lv, M.ColumnJoin):
          code, child = self.gen_cj_candidate(lv, rngs, "l1drv", "li", "h1", ind, "l1", "act", "k")
          o += code
          reg.update(child)
          bound.add(lv.var)
        else:
          o.append(f"{ind}h1 = act;")
          o += self.gen_cart(lv, rngs, "li", "act", ind, "a")
          for j, vs in enumerate(lv.var_from_source):
            bound.update(vs)
    for i, f in self.filters_ready(bound, done):
      o.append(f"{ind}h1 = h1 && {self.filter_code(f)};")
      done.add(i)
    if not deep:
      if len(done) != len(self.filters):
        raise CodegenError(f"{self.ep.rule}: filter over unbound variables")
      o.append(self.emit_code("h1", ind))
    else:
      self.gen_nested(0 if scan_root else 1, reg, bound, done, o, ind, "h1")
    o += ["    }",
          "    if (MODE == 2) {",
          "      srdj::seg_close(A, slice, seg_start, seg_used, seg_local);",
          "      if (seg_start != ~uint64_t(0)) { seg_start += seg_used; seg_room -= seg_used; }",
          "    }",
          "    if ((MODE == 0 || MODE == 2) && lane == 0) A.counts[slice] = cnt;",
          "  }"]
    if self.ring:
      o += ["  if (MODE == 2 && seg_start != ~uint64_t(0)) ring_out(seg_start, false);"]
    o += ["  if constexpr (MODE == 5) {",
          "    const uint64_t ps_ = srd::wave_sum<uint64_t>(uint64_t(probes_));",
          "    if (lane == 0 && ps_) atomicAdd(A.probe_ctr, (unsigned long long)ps_);",
          "  }",
          "}"]
    lds_bytes = 0
    for d in self._lds:
      dims = [int(x) for x in re.findall(r"\[(\d+)\]", d)]
      elem = 8 if "uint64_t" in d else (1 if "uint8_t" in d else 4)
      n = elem
      for x in dims:
        n *= x
      lds_bytes += n
    if lds_bytes > 64 * 1024:
      raise CodegenError(f"{self.ep.rule}: the join's level frames need {lds_bytes} bytes of LDS "
                         "per block (limit 65536); split the rule")
    # the ring only where eight blocks per CU still fit the CU's 160 KiB of LDS
    if (not self.ring and self.ring_ok and not self.direct_ok()
        and 8 * (lds_bytes + 4 * ring_cols * 128 * 4) <= 160 * 1024):
      self.ring = True
      try:
        return self.gen_run()
      except Exception:
        self.ring = False
        raise
    o[lds_at:lds_at] = self._lds
    return "\n".join(o)

  def needed_after(self, li: int) -> set[str]:
    """Variables read from level li on: source prefixes, checks, heads."""
    need: set[str] = set()
    for lv in self.levels[li:]:
      for s_ in lv.sources:
        need.update(s_.prefix)
    for f in self.filters:
      need.update(f.prefix if isinstance(f, M.Negation) else f.vars)
    for d in self.dests:
      need.update(d.terms)
    return need

  def gen_nested(self, li: int, reg: dict, bound: set[str], done: set[int], o: list[str], ind: str,
                 hvar: str) -> None:
    """Level li >= 1 of the join, emitted in wave-uniform control flow.

    Every lane holds one binding of the levels above (alive when `hvar`).
    The lanes publish their bindings and this level's source ranges in LDS
    frames, and the wave walks the concatenation of their iteration spaces
    64 items at a time; each item finds its owner lane through the 65-entry
    prefix array and re-reads the owner's frame.  Deeper levels repeat the
    pattern inside the batch loop, so every loop stays wave-uniform and the
    ballot-based emission at the innermost level stays valid."""
    if li == len(self.levels):
      if len(done) != len(self.filters):
        raise CodegenError(f"{self.ep.rule}: filter over unbound variables")
      o.append(self.emit_code(hvar, ind))
      return
    lv = self.levels[li]
    t = f"{li}"
    rngs = [self.resolve(reg, self.vid(s_), tuple(s_.prefix), o, ind, f"b{t}") for s_ in lv.sources]
    keep = sorted(v for v in bound if v in self.needed_after(li))
    frame = [self.var(v) for v in keep]
    per_lane = []
    for r in rngs:
      if r.full0:
        per_lane.append(None)
      else:
        per_lane.append((len(frame), len(frame) + 1))
        frame += [r.b, r.e]
    bm_at = {}
    bm_code = []  # emitted after the weigh mode's exit: weighing needs no bitmap slots
    for j, r in enumerate(rngs):
      if li == 1 and self.bm_eligible(r):  # level 2 only: deeper frames would outgrow LDS
        bm_code.append(f"{ind}const uint32_t hb{t}_{j} = ({hvar} && V.v[{r.view}].n_hv && "
                       f"{r.e} - {r.b} >= V.v[{r.view}].bm_min) ? "
                       f"srdj::heavy_slot(V.v[{r.view}], {r.b}) : ~0u;")
        bm_at[j] = len(frame)
        frame.append(f"hb{t}_{j}")
    nfr = len(frame)
    self._lds += [f"  __shared__ uint64_t pref_s{t}[4][65];",
                  f"  __shared__ uint32_t fr_s{t}[4][{nfr + 1}][64];",
                  f"  __shared__ uint32_t own_s{t}[4][64];",
                  f"  __shared__ uint8_t flg_s{t}[4][64];"]
    o.append(f"{ind}uint64_t sz{t} = 0;")
    o.append(f"{ind}uint32_t dv{t} = 0;")
    o.append(f"{ind}if ({hvar}) {{")
    if isinstance(lv, M.ColumnJoin):
      o.append(f"{ind}  sz{t} = {self.size_expr(rngs[0], False)};")
      for j in range(1, len(rngs)):
        o.append(f"{ind}  {{ const uint64_t z = {self.size_expr(rngs[j], False)}; "
                 f"if (z < sz{t}) {{ sz{t} = z; dv{t} = {j}; }} }}")
    else:
      o.append(f"{ind}  sz{t} = 1;")
      for r in rngs:
        o.append(f"{ind}  sz{t} *= uint64_t({self.size_expr(r, True)});")
    o.append(f"{ind}}}")
    if (isinstance(lv, M.ColumnJoin) and len(rngs) > 1 and not self.coop_search
        and not any(r.full0 for r in rngs)):
      self.gen_window(rngs, per_lane, frame, t, hvar, o, ind)
    if li == 1:
      o.append(f"{ind}if constexpr (MODE == 4) {{  // weigh: the item's level-2 candidates")
      o.append(f"{ind}  if (act) A.item_cost[item] = sz{t} + SRD_ITEM_COST;")
      o.append(f"{ind}  continue;")
      o.append(f"{ind}}}")
    tiled = (self.cart_tiles and isinstance(lv, M.CartesianJoin) and li == len(self.levels) - 1
             and len(rngs) == 2 and all(len(vs) == 1 for vs in lv.var_from_source))
    if tiled:
      # a large product leaves the flattened walk for the register-tiled one
      # below (its weigh cost above still counts it)
      sizes = [self.size_expr(r, True) for r in rngs]
      o.append(f"{ind}const bool big{t} = {hvar} && sz{t} >= SRD_CART_TILE_MIN && "
               f"max(uint32_t({sizes[0]}), uint32_t({sizes[1]})) >= SRD_CART_TILE_DIM;")
      o.append(f"{ind}if (big{t}) sz{t} = 0;")
    o += bm_code
    o.append(f"{ind}const uint64_t inc{t} = srd::wave_incl_scan<uint64_t>(sz{t});")
    o.append(f"{ind}const uint64_t tot{t} = srdj::uniform_from(inc{t}, 63);")
    o.append(f"{ind}__builtin_amdgcn_wave_barrier();")
    o.append(f"{ind}pref_s{t}[wv][lane + 1] = inc{t};")
    o.append(f"{ind}if (lane == 0) pref_s{t}[wv][0] = 0;")
    for fi, expr in enumerate(frame):
      o.append(f"{ind}fr_s{t}[wv][{fi}][lane] = {expr};")
    o.append(f"{ind}fr_s{t}[wv][{nfr}][lane] = dv{t};")
    # Owner lookup: the lanes with a non-empty space, in order, are listed in
    # own_s; per batch of 64 positions each such lane that starts inside the
    # batch flags its start, and a position's owner is the one of rank
    # (non-empty lanes starting before the batch) + (flags at or below it).
    o.append(f"{ind}const uint64_t ne{t} = srd::ballot(sz{t} > 0);")
    o.append(f"{ind}if (sz{t} > 0) own_s{t}[wv][srd::popc_below(ne{t})] = lane;")
    o.append(f"{ind}flg_s{t}[wv][lane] = 0;")
    o.append(f"{ind}const uint64_t st{t} = inc{t} - sz{t};")
    o.append(f"{ind}uint32_t base{t} = 0;")
    o.append(f"{ind}__builtin_amdgcn_wave_barrier();")
    o.append(f"{ind}__builtin_amdgcn_fence(__ATOMIC_SEQ_CST, \"wavefront\");")
    o.append(f"{ind}for (uint64_t j{t} = 0; j{t} < tot{t}; j{t} += 64) {{")
    i2 = ind + "  "
    o.append(f"{i2}if (sz{t} > 0 && st{t} >= j{t} && st{t} < j{t} + 64) flg_s{t}[wv][uint32_t(st{t} - j{t})] = 1;")
    o.append(f"{i2}__builtin_amdgcn_wave_barrier();")
    o.append(f"{i2}__builtin_amdgcn_fence(__ATOMIC_SEQ_CST, \"wavefront\");")
    o.append(f"{i2}const uint64_t sm{t} = srd::ballot(flg_s{t}[wv][lane] != 0);")
    o.append(f"{i2}flg_s{t}[wv][lane] = 0;")
    o.append(f"{i2}const uint32_t rk{t} = base{t} + srd::popc_below(sm{t}) + uint32_t((sm{t} >> lane) & 1ull);")
    o.append(f"{i2}base{t} += uint32_t(__popcll(sm{t}));")
    o.append(f"{i2}const uint64_t jj{t} = j{t} + lane;")
    o.append(f"{i2}const bool act{t} = jj{t} < tot{t};")
    o.append(f"{i2}const uint32_t own{t} = act{t} ? own_s{t}[wv][rk{t} - 1] : 0u;")
    o.append(f"{i2}const uint64_t loc{t} = act{t} ? jj{t} - pref_s{t}[wv][own{t}] : 0;")
    o.append(f"{i2}{{")
    i3 = i2 + "  "
    for fi, v in enumerate(keep):
      o.append(f"{i3}const uint32_t {self.var(v)} = fr_s{t}[wv][{fi}][own{t}];")
    reg2: dict = {k: r for k, r in reg.items() if r.uniform}
    rngs_o = []
    for j, r in enumerate(rngs):
      if per_lane[j] is None:
        rngs_o.append(r)
      else:
        bi, ei = per_lane[j]
        o.append(f"{i3}const uint32_t f{t}_{j}b = fr_s{t}[wv][{bi}][own{t}];")
        o.append(f"{i3}const uint32_t f{t}_{j}e = fr_s{t}[wv][{ei}][own{t}];")
        bm = ""
        if j in bm_at:
          o.append(f"{i3}const uint32_t f{t}_{j}h = fr_s{t}[wv][{bm_at[j]}][own{t}];")
          bm = f"f{t}_{j}h"
        rr = Rg(r.view, r.depth, f"f{t}_{j}b", f"f{t}_{j}e", bm=bm)
        rngs_o.append(rr)
        s_ = lv.sources[j]
        reg2[(self.vid(s_), tuple(s_.prefix))] = rr
    o.append(f"{i3}const uint32_t dvo{t} = fr_s{t}[wv][{nfr}][own{t}];")
    hn = f"h{li + 1}"
    o.append(f"{i3}bool {hn} = false;")
    bound2 = set(keep)
    if isinstance(lv, M.ColumnJoin):
      code, child = self.gen_cj_candidate(lv, rngs_o, f"dvo{t}", f"loc{t}", hn, i3, f"l{li + 1}",
                                          f"act{t}", f"own{t}")
      o += code
      reg2.update(child)
      bound2.add(lv.var)
    else:
      o.append(f"{i3}{hn} = act{t};")
      o += self.gen_cart(lv, rngs_o, f"loc{t}", f"act{t}", i3, f"c{t}")
      for vs in lv.var_from_source:
        bound2.update(vs)
    done2 = set(done)
    for i, f in self.filters_ready(bound2, done2):
      o.append(f"{i3}{hn} = {hn} && {self.filter_code(f)};")
      done2.add(i)
    self.gen_nested(li + 1, reg2, bound2, done2, o, i3, hn)
    o.append(f"{i2}}}")
    o.append(f"{i2}__builtin_amdgcn_wave_barrier();")
    o.append(f"{ind}}}")
    if tiled:
      self.gen_cart_tiles(lv, li, rngs, per_lane, keep, done, o, ind)

  def gen_window(self, rngs: list[Rg], per_lane: list, frame: list, t: str, hvar: str,
                 o: list[str], ind: str) -> None:
    """Narrow each item's driving range to the values every other source
    can hold: a value of the intersection is at least the largest first value
    and at most the smallest last value of the sources' ranges (each range
    is sorted), so the driver is cut to [lower_bound(largest first),
    upper_bound(smallest last)) -- the first seek of the reference's
    leapfrog intersection (advance_to_intersection,
    runtime/generalized_datalog/gpu/device_sorted_array_index.ipp:1390-1433),
    taken once per item instead of per candidate.  Items with fewer than
    A.window_min candidates keep their range (two searches would cost more
    than they save), and so does the weigh mode: an item's cost stays its
    whole driving range, which spares the weigh pass the searches (on the
    R-MAT scale-20 triangle join they tripled it, 0.37 -> 1.07 ms).  The
    frame then carries the window as the driver's range; the other sources'
    ranges, their bitmap slots and every search are unchanged."""
    cols = [self.col_expr(r.view, r.depth) for r in rngs]

    def pick(vals):
      e = vals[-1]
      for j in range(len(vals) - 2, -1, -1):
        e = f"(dv{t} == {j}u ? {vals[j]} : {e})"
      return e
    i1, i2 = ind + "  ", ind + "    "
    o.append(f"{ind}uint32_t wb{t} = 0, we{t} = 0;  // the driver's window")
    o.append(f"{ind}if ({hvar}) {{")
    o.append(f"{i1}wb{t} = {pick([f'uint32_t({r.b})' for r in rngs])};")
    o.append(f"{i1}we{t} = {pick([f'uint32_t({r.e})' for r in rngs])};")
    o.append(f"{i1}if (MODE != 4 && A.window_min && sz{t} >= A.window_min) {{")
    o.append(f"{i2}uint32_t lv_ = {cols[0]}[{rngs[0].b}], hv_ = {cols[0]}[{rngs[0].e} - 1];")
    for j in range(1, len(rngs)):
      o.append(f"{i2}lv_ = max(lv_, {cols[j]}[{rngs[j].b}]); "
               f"hv_ = min(hv_, {cols[j]}[{rngs[j].e} - 1]);")
    o.append(f"{i2}const uint32_t* cd_ = {pick(cols)};")
    o.append(f"{i2}wb{t} = srd::lower_bound_la(cd_, wb{t}, we{t}, lv_);")
    o.append(f"{i2}we{t} = lv_ <= hv_ ? srd::upper_bound_la(cd_, wb{t}, we{t}, hv_) : wb{t};")
    o.append(f"{i2}sz{t} = we{t} - wb{t};")
    o.append(f"{i2}if constexpr (MODE == 5) probes_ += 2u;")
    o.append(f"{i1}}}")
    o.append(f"{ind}}}")
    for j, r in enumerate(rngs):
      bi, ei = per_lane[j]
      frame[bi] = f"(dv{t} == {j}u ? wb{t} : uint32_t({r.b}))"
      frame[ei] = f"(dv{t} == {j}u ? we{t} : uint32_t({r.e}))"

  def gen_cart_tiles(self, lv: M.CartesianJoin, li: int, rngs: list[Rg], per_lane: list, keep: list,
                     done: set[int], o: list[str], ind: str) -> None:
    """Register-tiled walk of the large two-source cartesian products of a
    nested last level (the reference stages 128-row tiles of both sources
    in shared memory, ws_infrastructure.h:499-604).  The owners, one at a
    time, wave-uniform: the lanes hold 64 rows of the longer source, a
    64-row tile of the other is loaded one row per lane and broadcast a row
    at a time by readlane, so every step emits up to 64 tuples with no
    per-tuple index arithmetic.  The tuples are the product's, in another
    order than the flattened walk's; count and write passes run the same
    code, so their slice counts and offsets agree."""
    t = f"{li}"
    i1 = ind + "  "
    i2 = i1 + "  "
    o.append(f"{ind}uint64_t bm{t} = srd::ballot(big{t});")
    o.append(f"{ind}while (bm{t}) {{")
    o.append(f"{i1}const uint32_t ow{t} = uint32_t(__builtin_ctzll(bm{t}));")
    o.append(f"{i1}bm{t} &= bm{t} - 1;")
    # the owner's frame, read by every lane and marked wave-uniform so the
    # tile loops run on scalar counters
    rfl = "__builtin_amdgcn_readfirstlane"
    for fi, v in enumerate(keep):
      o.append(f"{i1}const uint32_t {self.var(v)} = {rfl}(fr_s{t}[wv][{fi}][ow{t}]);")
    be = []
    for j, r in enumerate(rngs):
      if per_lane[j] is None:
        be.append((r.b, r.e))
      else:
        bi, ei = per_lane[j]
        o.append(f"{i1}const uint32_t tb{t}_{j} = {rfl}(fr_s{t}[wv][{bi}][ow{t}]);")
        o.append(f"{i1}const uint32_t te{t}_{j} = {rfl}(fr_s{t}[wv][{ei}][ow{t}]);")
        be.append((f"tb{t}_{j}", f"te{t}_{j}"))
    cols = [self.col_expr(r.view, r.depth) for r in rngs]
    o.append(f"{i1}const uint32_t n{t}_0 = {be[0][1]} - {be[0][0]}, n{t}_1 = {be[1][1]} - {be[1][0]};")
    o.append(f"{i1}const bool l1{t} = n{t}_1 >= n{t}_0;  // the lanes walk the longer source")
    o.append(f"{i1}const uint32_t nl{t} = l1{t} ? n{t}_1 : n{t}_0, nr{t} = l1{t} ? n{t}_0 : n{t}_1;")
    o.append(f"{i1}const uint32_t* cl{t} = l1{t} ? {cols[1]} + {be[1][0]} : {cols[0]} + {be[0][0]};")
    o.append(f"{i1}const uint32_t* cr{t} = l1{t} ? {cols[0]} + {be[0][0]} : {cols[1]} + {be[1][0]};")
    o.append(f"{i1}for (uint32_t tl{t} = 0; tl{t} < nl{t}; tl{t} += 64) {{")
    o.append(f"{i2}const bool inl{t} = tl{t} + lane < nl{t};")
    o.append(f"{i2}const uint32_t wl{t} = inl{t} ? cl{t}[tl{t} + lane] : 0u;")
    o.append(f"{i2}for (uint32_t tr{t} = 0; tr{t} < nr{t}; tr{t} += 64) {{")
    i3 = i2 + "  "
    o.append(f"{i3}const uint32_t cn{t} = min(64u, nr{t} - tr{t});")
    o.append(f"{i3}const uint32_t wr{t} = lane < cn{t} ? cr{t}[tr{t} + lane] : 0u;")
    o.append(f"{i3}for (uint32_t a{t} = 0; a{t} < cn{t}; ++a{t}) {{")
    i4 = i3 + "  "
    o.append(f"{i4}const uint32_t vr{t} = __builtin_amdgcn_readlane(wr{t}, a{t});")
    (va,), (vb,) = lv.var_from_source
    o.append(f"{i4}const uint32_t {self.var(va)} = l1{t} ? vr{t} : wl{t};")
    o.append(f"{i4}const uint32_t {self.var(vb)} = l1{t} ? wl{t} : vr{t};")
    hn = f"ht{t}"
    o.append(f"{i4}bool {hn} = inl{t};")
    bound2 = set(keep) | {va, vb}
    done2 = set(done)
    for i, f in self.filters_ready(bound2, done2):
      o.append(f"{i4}{hn} = {hn} && {self.filter_code(f)};")
      done2.add(i)
    if len(done2) != len(self.filters):
      raise CodegenError(f"{self.ep.rule}: filter over unbound variables")
    o.append(self.emit_code(hn, i4))
    o.append(f"{i3}}}")
    o.append(f"{i2}}}")
    o.append(f"{i1}}}")
    o.append(f"{ind}}}")

  # --------------------------------------------------------- assembly ----
  def gen_kernels(self) -> str:
    out = [f"namespace {self.ns} {{", self.gen_root(), "", self.gen_run(), "",
           "void launch_root(const SrdLaunchArgs* a, hipStream_t s) {",
           "  root_k<false><<<(a->n_root + 255u) / 256u, 256, 0, s>>>(*a);",
           "}",
           "void launch_count(const SrdLaunchArgs* a, hipStream_t s) {",
           "  run_k<0, false><<<a->grid_blocks, 256, 0, s>>>(*a);",
           "}",
           "void launch_write(const SrdLaunchArgs* a, hipStream_t s) {",
           "  run_k<1, false><<<a->grid_blocks, 256, 0, s>>>(*a);",
           "}",
           "void launch_fused(const SrdLaunchArgs* a, hipStream_t s) {",
           "  run_k<2, false><<<a->grid_blocks, 256, 0, s>>>(*a);",
           "}",
           "void launch_direct(const SrdLaunchArgs* a, hipStream_t s) {",
           "  run_k<3, false><<<a->grid_blocks, 256, 0, s>>>(*a);",
           "}",
           "void launch_probe(const SrdLaunchArgs* a, hipStream_t s) {",
           "  run_k<5, false><<<a->grid_blocks, 256, 0, s>>>(*a);",
           "}"] + ([
           "void launch_mark(const SrdLaunchArgs* a, hipStream_t s) {",
           "  run_k<6, false><<<a->grid_blocks, 256, 0, s>>>(*a);",
           "}"] if self.mark_ok() else []) + ([
           "void launch_weigh(const SrdLaunchArgs* a, hipStream_t s) {",
           "  run_k<4, false><<<a->grid_blocks, 256, 0, s>>>(*a);",
           "}"] if len(self.levels) >= 2 else []) + ([
           "// device-sized launches (a->dyn holds views and sizes; fixed grids)",
           "void launch_root_dyn(const SrdLaunchArgs* a, hipStream_t s) {",
           "  root_k<true><<<a->grid_blocks, 256, 0, s>>>(*a);",
           "}",
           "void launch_count_dyn(const SrdLaunchArgs* a, hipStream_t s) {",
           "  run_k<0, true><<<a->grid_blocks, 256, 0, s>>>(*a);",
           "}",
           "void launch_write_dyn(const SrdLaunchArgs* a, hipStream_t s) {",
           "  run_k<1, true><<<a->grid_blocks, 256, 0, s>>>(*a);",
           "}"] + ([
           "void launch_direct_dyn(const SrdLaunchArgs* a, hipStream_t s) {",
           "  run_k<3, true><<<a->grid_blocks, 256, 0, s>>>(*a);",
           "}"] if self.direct_ok() else []) if self.dyn else []) + [
           f"}}  // namespace {self.ns}", ""]
    return "\n".join(out)

  def gen_desc(self) -> str:
    views = []
    bmv = self.bm_views()
    rpv = self.root_pos_views() if not self.coop_search else set()
    for vi, (rel, index, ver) in enumerate(self.info.views):
      cols = ", ".join(str(c) for c in index)
      views.append(f"{{\"{rel}\", {len(index)}, {{{cols}}}, {VER_CODE[ver]}, "
                   f"{1 if vi in bmv else 0}, {1 if vi in rpv else 0}}}")
    if isinstance(self.root, M.Scan):
      kind, roots = "SRD_ROOT_SCAN", [self.vid(self.root)]
    else:
      kind, roots = "SRD_ROOT_JOIN", [self.vid(s) for s in self.root.sources]
    dests = []
    for d in self.dests:
      cols = ", ".join(str(c) for c in d.index)
      dests.append(f"{{\"{d.rel}\", {len(d.index)}, {{{cols}}}}}")
    grain = 256 if len(self.levels) == 2 else 4096
    return ("  {" + f"\"{self.ep.rule}\", {len(views)}, {{{', '.join(views)}}}, {kind}, "
            f"{len(roots)}, {{{', '.join(str(r) for r in roots)}}}, {len(dests)}, "
            f"{{{', '.join(dests)}}}, {1 if self.ep.count else 0}, {grain}, "
            f"{1 if self.direct_ok() else 0}, "
            f"&{self.ns}::launch_root, &{self.ns}::launch_count, &{self.ns}::launch_write, "
            f"&{self.ns}::launch_fused, &{self.ns}::launch_direct, "
            f"{'&' + self.ns + '::launch_weigh' if len(self.levels) >= 2 else 'nullptr'}, "
            f"{1 if self.sorted_output() else 0}, &{self.ns}::launch_probe, "
            + (f"&{self.ns}::launch_root_dyn, &{self.ns}::launch_count_dyn, "
               f"&{self.ns}::launch_write_dyn, "
               + (f"&{self.ns}::launch_direct_dyn" if self.direct_ok() else "nullptr")
               if self.dyn else "nullptr, nullptr, nullptr, nullptr")
            + (f", &{self.ns}::launch_mark" if self.mark_ok() else ", nullptr") + "}")


@dataclass
class RelDecl:
  name: str
  arity: int
  input_file: str = ""
  print_size: bool = False


def program_text(decls: list[RelDecl], prog: M.Program) -> str:
  lines = ["(relations"]
  for d in decls:
    s = f"  (relation :name {d.name} :arity {d.arity}"
    if d.input_file:
      s += f' :input-file "{d.input_file}"'
    if d.print_size:
      s += " :print-size true"
    lines.append(s + ")")
  lines.append(")")
  return "\n".join(lines) + "\n" + M.print_program(prog) + "\n"


SHIM = r"""
// ======== extern "C" shim (same symbols as the reference's generated shim) ========
static srd_engine* g_eng = nullptr;

static bool ensure_engine() {
  if (!g_eng) g_eng = srd_engine_create(k_program, k_pipes, sizeof(k_pipes) / sizeof(k_pipes[0]));
  return g_eng && srd_engine_ok(g_eng);
}

extern "C" {

int srdatalog_init(void) {
  try {
    return ensure_engine() ? 0 : 1;
  } catch (...) {
    return 2;
  }
}

int srdatalog_load_csv(const char* rel_name, const char* path) {
  if (!rel_name || !path) return 1;
  if (!ensure_engine()) return 3;
  return srd_engine_load_csv(g_eng, rel_name, path);
}

int srdatalog_load_all(const char* data_dir) {
  if (!data_dir) return 1;
  if (!ensure_engine()) return 2;
  return srd_engine_load_all(g_eng, data_dir);
}

int srdatalog_run(unsigned long long max_iters) {
  if (!g_eng) return 1;
  return srd_engine_run(g_eng, max_iters);
}

unsigned long long srdatalog_size(const char* rel_name) {
  if (!g_eng || !rel_name) return 0;
  return srd_engine_size(g_eng, rel_name);
}

int srdatalog_shutdown(void) {
  if (g_eng) {
    srd_engine_destroy(g_eng);
    g_eng = nullptr;
  }
  return 0;
}

int srdatalog_export(const char* rel_name, uint32_t* out, uint64_t cap_rows, uint64_t* n_rows) {
  if (!rel_name || !n_rows) return 1;
  if (!g_eng) return 2;
  return srd_engine_export(g_eng, rel_name, out, cap_rows, n_rows);
}

int srdatalog_delta_sizes(const char* rel_name, unsigned long long* out, unsigned long long cap) {
  if (!g_eng || !rel_name) return -1;
  return srd_engine_delta_sizes(g_eng, rel_name, out, cap);
}

unsigned long long srdatalog_num_iterations(const char* rel_name) {
  if (!g_eng || !rel_name) return 0;
  return srd_engine_num_iterations(g_eng, rel_name);
}

int srdatalog_load_rows(const char* rel_name, const uint32_t* rows, uint64_t n, uint32_t arity) {
  if (!rel_name) return 1;
  if (!ensure_engine()) return 3;
  return srd_engine_load_rows(g_eng, rel_name, rows, n, arity);
}

int srdatalog_prepare(void) {
  if (!ensure_engine()) return 2;
  return srd_engine_prepare(g_eng);
}

double srdatalog_run_seconds(void) { return g_eng ? srd_engine_run_seconds(g_eng) : 0.0; }

unsigned long long srdatalog_join_tuples(void) { return g_eng ? srd_engine_join_tuples(g_eng) : 0; }

void srdatalog_count_probes(int on) {
  if (ensure_engine()) srd_engine_count_probes(g_eng, on);
}

unsigned long long srdatalog_join_probes(void) { return g_eng ? srd_engine_join_searches(g_eng) : 0; }

void srdatalog_set_grain(unsigned grain) {
  if (ensure_engine()) srd_engine_set_grain(g_eng, grain);
}

int srdatalog_set_instances(const uint32_t* bounds, uint64_t n_bounds) {
  if (!ensure_engine()) return 2;
  return srd_engine_set_instances(g_eng, bounds, n_bounds);
}

long long srdatalog_instance_delta_sizes(const char* rel_name, unsigned long long* out,
                                         unsigned long long cap_iters) {
  if (!g_eng) return -1;
  return srd_engine_instance_delta_sizes(g_eng, rel_name, out, cap_iters);
}

int srdatalog_export_instances(const char* rel_name, uint32_t* out, uint64_t cap_rows,
                               uint64_t* n_rows, uint64_t* cuts) {
  if (!g_eng) return 3;
  return srd_engine_export_instances(g_eng, rel_name, out, cap_rows, n_rows, cuts);
}

// ---- engine handles: resumable runs for iteration-granular serving ----
void* srdatalog_engine_new(void) {
  try {
    srd_engine* e = srd_engine_create(k_program, k_pipes, sizeof(k_pipes) / sizeof(k_pipes[0]));
    if (e && !srd_engine_ok(e)) {
      srd_engine_destroy(e);
      return nullptr;
    }
    return e;
  } catch (...) {
    return nullptr;
  }
}

void srdatalog_engine_free(void* h) {
  if (h) srd_engine_destroy(static_cast<srd_engine*>(h));
}

void srdatalog_engine_set_quiet(void* h, int quiet) {
  if (h) srd_engine_set_quiet(static_cast<srd_engine*>(h), quiet);
}

int srdatalog_engine_load_rows(void* h, const char* rel_name, const uint32_t* rows, uint64_t n,
                               uint32_t arity) {
  if (!h || !rel_name) return 1;
  return srd_engine_load_rows(static_cast<srd_engine*>(h), rel_name, rows, n, arity);
}

int srdatalog_engine_run_begin(void* h) {
  if (!h) return 1;
  return srd_engine_run_begin(static_cast<srd_engine*>(h));
}

int srdatalog_engine_run_iterations(void* h, unsigned long long k, unsigned long long* done,
                                    int* converged) {
  if (!h) return 1;
  return srd_engine_run_iterations(static_cast<srd_engine*>(h), k, done, converged);
}

int srdatalog_engine_absorb(void* dst, void* src) {
  if (!dst || !src) return 1;
  return srd_engine_absorb(static_cast<srd_engine*>(dst), static_cast<srd_engine*>(src));
}

int srdatalog_engine_range_counts(void* h, const char* rel_name, int delta, const uint32_t* lohi,
                                  uint32_t m, unsigned long long* out) {
  if (!h) return 1;
  return srd_engine_range_counts(static_cast<srd_engine*>(h), rel_name, delta, lohi, m, out);
}

int srdatalog_engine_export_ranges(void* h, const char* rel_name, const uint32_t* lohi, uint32_t m,
                                   int rebase, uint32_t* out, uint64_t cap_rows, uint64_t* n_rows,
                                   uint64_t* cuts) {
  if (!h) return 1;
  return srd_engine_export_ranges(static_cast<srd_engine*>(h), rel_name, lohi, m, rebase, out,
                                  cap_rows, n_rows, cuts);
}

int srdatalog_engine_erase_ranges(void* h, const uint32_t* lohi, uint32_t m) {
  if (!h) return 1;
  return srd_engine_erase_ranges(static_cast<srd_engine*>(h), lohi, m);
}

}  // extern "C"
"""


def generate_program_source(name: str, decls: list[RelDecl], prog: M.Program) -> str:
  """Complete HIP translation unit of a program."""
  arity = {d.name: d.arity for d in decls}
  rec = set()
  for st in prog.steps:
    if st.recursive and isinstance(st.body, M.FixpointPlan):
      for op in st.body.ops:
        for ep in (op.ops if isinstance(op, M.ParallelGroup) else [op]):
          if isinstance(ep, M.ExecutePipeline):
            rec.add(id(ep))
  # SRDATALOG_CART_TILES=0 (at generation time): every cartesian level takes
  # the flattened per-tuple walk
  cart = os.environ.get("SRDATALOG_CART_TILES", "1") != "0"
  gens = [PipelineGen(ep, arity, i, dyn=id(ep) in rec, cart_tiles=cart)
          for i, ep in enumerate(prog.pipelines())]
  seen = set()
  for g in gens:
    if g.ep.rule in seen:
      raise CodegenError(f"duplicate pipeline rule name {g.ep.rule}")
    seen.add(g.ep.rule)
  text = program_text(decls, prog)
  if ")SRDMIR\"" in text:
    raise CodegenError("program text contains the raw-string delimiter")
  out = [f"// Generated by srdatalog_hip.codegen for program {name}. Do not edit.",
         "#include <hip/hip_runtime.h>",
         "#include <stdint.h>",
         '#include "srd_join.h"',
         '#include "srd_engine.h"',
         '#include "srd_pipeline.h"',
         ""]
  for g in gens:
    out.append(g.gen_kernels())
  out.append("static const SrdPipelineDesc k_pipes[] = {")
  out.append(",\n".join(g.gen_desc() for g in gens) if gens else "")
  out.append("};")
  out.append(f'static const char* k_program = R"SRDMIR({text})SRDMIR";')
  out.append(SHIM)
  return "\n".join(out)
