"""Static properties the join code generator derives from a rule's plan:
which rules write their rows already sorted and duplicate-free (the engine
then adopts NEW as an index without checking it) and which views get
heavy-key bitmaps.  A wrong "sorted" claim would corrupt NEW, so the claim
is checked against plans where it must and must not hold."""

from srdatalog_hip import codegen, compiler, programs


def gens(prog):
  decls = compiler.relation_decls(prog)
  arity = {d.name: d.arity for d in decls}
  mir = compiler.compile_to_mir(prog)
  return {ep.rule: codegen.PipelineGen(ep, arity, i) for i, ep in enumerate(mir.pipelines())}


def test_triangle_join_is_sorted_output():
  g = gens(programs.PROGRAMS["triangle"]())
  assert g["TriangleJoin"].sorted_output()
  assert not g["KnowsLoad"].sorted_output()          # a scan root
  assert g["TriangleJoin"].bm_views() == {0}          # Knows, searched at depth 1 by level 2
  assert g["TriangleJoin"].root_pos_views() == {0}    # Knows(y, _): level 1 looks y up
  assert g["KnowsLoad"].root_pos_views() == set()


def test_rules_with_permuted_or_projected_heads_are_not_sorted_output():
  # none of these writes (root var, level-1 var, ...) in its destination's
  # index order: TC's recursive rule is rooted at the join variable y, which
  # its head drops; the others project join variables away or reorder them
  for name in ("tc", "sg", "andersen", "q3", "q7", "neg2hop", "reach"):
    for rule, g in gens(programs.PROGRAMS[name]()).items():
      assert not g.sorted_output(), (name, rule)


def test_generated_desc_carries_the_flags():
  src = codegen.generate_program_source("triangle", compiler.relation_decls(programs.PROGRAMS["triangle"]()),
                                        compiler.compile_to_mir(programs.PROGRAMS["triangle"]()))
  line = next(l for l in src.splitlines() if l.lstrip().startswith('{"TriangleJoin"'))
  # weigh launcher, sorted_output = 1, probe launcher; a non-recursive rule
  # has no device-sized launches
  assert line.rstrip().endswith("launch_weigh, 1, &srdp_1_TriangleJoin::launch_probe, "
                                "nullptr, nullptr, nullptr, nullptr, &srdp_1_TriangleJoin::launch_mark}")
  # want_bitmaps, want_root_pos
  assert '{"Knows", 2, {0, 1}, SRD_FULL, 1, 1}' in line


def test_recursive_rules_get_device_sized_launches():
  """Rules of recursive steps carry root/count/write (and, with one tuple
  per item, direct) launches that read their sizes from the device."""
  prog = programs.PROGRAMS["tc"]()
  src = codegen.generate_program_source("tc", compiler.relation_decls(prog),
                                        compiler.compile_to_mir(prog))
  rec = next(l for l in src.splitlines() if l.lstrip().startswith('{"TCRec'))
  assert rec.rstrip().rstrip(",").endswith("::launch_mark}")
  assert "::launch_direct_dyn, &srdp_" in rec
  assert "launch_root_dyn" in rec and "launch_count_dyn" in rec and "launch_write_dyn" in rec
  base = next(l for l in src.splitlines() if l.lstrip().startswith('{"TCBase'))
  assert base.rstrip().rstrip(",").endswith("nullptr, nullptr, nullptr, nullptr, &srdp_1_TCBase::launch_mark}")


def test_mark_mode_needs_one_all_variable_destination():
  """Mode 6 (dense-bitmap NEW) is generated for rules with one destination
  whose terms are all variables, keyed in the destination's index order."""
  prog = programs.PROGRAMS["andersen"]()
  gs = gens(prog)
  assert all(g.mark_ok() for g in gs.values())
  src = codegen.generate_program_source("andersen", compiler.relation_decls(prog),
                                        compiler.compile_to_mir(prog))
  assert "run_k<6, false>" in src and "A.mark_bits + (key_ >> 5)" in src
  # count-only rules write nothing
  cnt = gens(programs.PROGRAMS["triangle_count"]())
  assert not cnt["TriangleJoin"].mark_ok()
  assert "run_k<3, true>" in src and "root_k<true>" in src
