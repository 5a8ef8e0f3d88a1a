"""The pinning checks catch plausible mistakes.  Each planted mistake
(oracle/rc_oracle.c, orc_set_mutation) must make at least one of the
oracle's pinning checks fail: the golden vectors of the compiled reference,
the pregrid probe, the small-kernel vectors, the strict in-bound KAT, the
transpose identity, the Hogbom brute-force checks and the binner's
geometric, KAT and order checks."""
import pytest

import test_oracle_binner as TB
import test_oracle_golden as TG
import test_oracle_math as TM
from _fixtures import RANDOM_FIXTURES
from oracle import oracle as O

CHECKS = (
    [(f"golden {n}", lambda n=n: TG.test_random_fixture_bit_exact(n)) for n in RANDOM_FIXTURES]
    + [("pregrid probe", TG.test_pregrid_probe_matches_reference),
       ("small kernels", TG.test_small_kernels_bit_exact),
       ("strict in-bound KAT", TM.test_out_of_bound_strict_upper_limit),
       ("transpose identity", lambda: TM.test_degrid_is_the_transpose(True)),
       ("hogbom point source", TM.test_hogbom_recovers_point_source_exactly),
       ("hogbom negative peaks", TM.test_hogbom_residual_identity_and_negative_peaks),
       ("binner copies and order", TB.test_records_copies_and_order),
       ("binner remainder KATs", TB.test_remainder_edge_kats)]
)

MUTATIONS = {
    1: "transposed tap operand",
    2: "HalfGCF conjugation dropped",
    3: "oversampling offset rounded, not floored",
    4: "window recentring dropped",
    5: "degridder conjugates (adjoint, not transpose)",
    6: "in-bound test not strict",
    7: "w-plane from the scaled w",
    8: "normalisation by size^2, not size*pitch",
    9: "phase rotation sign flipped",
    10: "Hogbom peak by signed value",
    11: "binner right margin not strict",
    12: "binner floor division",
    13: "binner bin order broken",
}


def failures():
    out = []
    for name, fn in CHECKS:
        try:
            fn()
        except AssertionError:
            out.append(name)
    return out


def test_checks_pass_without_mutation():
    assert failures() == []


@pytest.mark.parametrize("m", sorted(MUTATIONS))
def test_planted_mistake_is_caught(m):
    h = O.oracle_lib()
    h.orc_set_mutation(m)
    try:
        caught = failures()
    finally:
        h.orc_set_mutation(0)
    assert caught, f"mutation {m} ({MUTATIONS[m]}) passed every pinning check"
