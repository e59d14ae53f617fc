"""The algorithmic operation counts that price a build (bench.py's roofline).

hfcode_amd/gen/flops_per_class.json is frozen; the generator's model (one Boys
evaluation and one vertical recursion per primitive quartet over the union of
a class's targets) must still reproduce it, and must never price more than
the emitted kernels execute."""
import json

import pytest

from hfcode_amd.gen import gen_eri


@pytest.fixture(scope="module")
def counts():
    return gen_eri.algorithmic_counts()


def test_frozen_table_matches_generator(counts):
    with open(gen_eri.FLOPS_FILE) as fh:
        frozen = json.load(fh)["classes"]
    assert sorted(frozen) == sorted(gen_eri.CLASSES)
    for cls in gen_eri.CLASSES:
        for k in gen_eri.FROZEN_KEYS:
            assert frozen[cls][k] == counts[cls][k], (cls, k)


def test_union_never_exceeds_executed(counts):
    for cls, c in counts.items():
        # per primitive quartet (the dominant term): one shared Boys evaluation and
        # recursion cannot cost more than the emitted kernels' repeated ones
        assert c["flops_prim"] <= c["exec_prim"], cls
        assert c["flops_quartet"] == c["exec_quartet"], cls


def test_ssss_closed_form(counts):
    """(ss|ss): per primitive quartet 24 set-up operations (1/(p+q), rho_p,
    rho_q, T, 1/2q, 1/2(p+q), W-P, W-Q, Q-C, P-Q), the Boys table evaluation of
    F_0 (7 Taylor factors + 15 for the series + 3 for the index), the
    prefactor (3), s0 = pref F_0 (1) and the accumulation (1): 54; per
    quartet 12 contraction operations + the block maximum: 13."""
    c = counts["ssss"]
    assert c["flops_prim"] == 24 + (7 + 15 + 3) + 3 + 1 + 1 == 54
    assert c["flops_quartet"] == 13
    assert not c["split"]


def test_dlll_counts():
    """(dl|ll), the dominant C60 class: one Boys evaluation (F_0..F_5) per
    primitive quartet instead of one per emitted kernel."""
    with open(gen_eri.FLOPS_FILE) as fh:
        c = json.load(fh)["classes"]["dlll"]
    boys = gen_eri.boys_flops(5, gen_eri.BOYS_TERMS["double"])
    assert boys == 100
    assert c["flops_prim"] < c["exec_prim"]
    assert c["flops_quartet"] == c["exec_quartet"] == 5968
