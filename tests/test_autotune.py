"""The grouping autotuner's selection (hfcode_amd/gen/autotune_groups.py):
per class and precision the budget of the fastest variant, and the tuned
table changes how the generator groups a class's combinations."""
import json

from hfcode_amd.gen import autotune_groups as T
from hfcode_amd.gen import gen_eri as G


def _fake_measure(path):
    classes = {c: [0.0, 0.0] for c in G.CLASSES}
    def cms(fp64_dlll, fp32_dlll, fp64_slll, fp32_slll):
        f = {c: list(v) for c, v in classes.items()}
        m = {c: list(v) for c, v in classes.items()}
        f["dlll"] = [fp64_dlll, 0.0]
        m["dlll"] = [1.0, fp32_dlll]
        f["slll"] = [fp64_slll, 0.0]
        m["slll"] = [1.0, fp32_slll]
        return {"fp64": f, "mixed": m}
    json.dump({"variants": {"b32": [32, 64], "b128": [128, 256]}, "workload": "test",
               "class_ms": {"b32": cms(10.0, 5.0, 3.0, 2.0), "b128": cms(8.0, 6.0, 4.0, 1.0)}},
              open(path, "w"))


def test_select_picks_fastest_budget_per_class_and_precision(tmp_path):
    mp, out = tmp_path / "m.json", tmp_path / "b.json"
    _fake_measure(mp)
    T.select(str(mp), str(out))
    b = json.load(open(out))["budgets"]
    assert b["dlll"] == [128, 64]       # FP64 faster with b128, FP32 with b32
    assert b["slll"] == [32, 256]
    assert "ssss" not in b              # fused classes have no grouping
    assert "dddd" not in b              # not measured (zero time): default budgets


def test_budgets_change_the_grouping(monkeypatch):
    g = G.Gen("llll")
    parts = len(g.bra_parts) * len(g.ket_parts)
    monkeypatch.setattr(G, "GROUP_BUDGETS", {"llll": (1, 1)})
    assert len(g.group_plan(True)) == parts                 # every combination alone
    monkeypatch.setattr(G, "GROUP_BUDGETS", {"llll": (10 ** 6, 10 ** 6)})
    assert len(g.group_plan(True)) == len(g.bra_parts)     # one kernel per bra part
    assert len(g.group_plan(True, True)) == len(g.bra_parts)
    for plan in (g.group_plan(True), g.group_plan(True, True)):
        assert sorted(c for grp in plan for c in grp) == sorted((bp, kp) for bp in g.bra_parts for kp in g.ket_parts)
