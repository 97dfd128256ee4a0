"""The oracle against the committed golden vectors of the reference's CPU
library (tests/golden/ref_vectors.npz, written by tests/golden/make_golden.py
from the reference's own sources). Needs no reference tree: the inputs are
regenerated from the same seeds and the oracle's outputs must hash to the
stored digests."""
import hashlib
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden  # noqa: E402
import oracle_lib as ol  # noqa: E402

FIX = os.path.join(HERE, "golden", "ref_vectors.npz")


class OracleAsRef:
    """The oracle's entry points under the names make_golden uses."""

    def __init__(self, o):
        self.o = o

    def __getattr__(self, name):
        base = name[len("orc_"):]
        return getattr(self.o, "orc_query" if base == "query_seeded" else "orc_" + base)


def test_oracle_matches_reference_golden_vectors():
    gold = np.load(FIX, allow_pickle=False)
    names = sorted({k.split("__")[0] for k in gold.files})
    seen = set()
    for name, arr in make_golden.cases(OracleAsRef(ol.oracle()), "orc_"):
        seen.add(name)
        assert np.array_equal(np.ascontiguousarray(arr[:8], np.uint64), gold[name + "__head"]), name
        got = np.frombuffer(hashlib.sha256(np.ascontiguousarray(arr).tobytes()).digest(), np.uint8)
        assert np.array_equal(got, gold[name + "__sha256"]), name
    assert seen == set(names), f"fixture/case mismatch: {set(names) ^ seen}"


@pytest.mark.skipif(not ol.have_ref(), reason="reference CPU library not built here")
def test_golden_vectors_are_current():
    """Regenerating from the reference library gives the committed digests."""
    gold = np.load(FIX, allow_pickle=False)
    for name, arr in make_golden.cases(ol.ref(), "ref_"):
        got = np.frombuffer(hashlib.sha256(np.ascontiguousarray(arr).tobytes()).digest(), np.uint8)
        assert np.array_equal(got, gold[name + "__sha256"]), name
