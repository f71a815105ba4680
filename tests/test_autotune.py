"""Autotuner bookkeeping (CPU): cache keys and the JSON cache."""
from pdekit_amd import _lib as L
from pdekit_amd import autotune as A
from pdekit_amd.mesh import rectangle


def test_cache_key_separates_kinds_and_sizes():
    m1 = rectangle(8, 8, split=2, order=2)
    m2 = rectangle(8, 8, split=2, order=1)
    m3 = rectangle(64, 64, split=2, order=2)
    k = lambda m, ph=L.BURGERS2D, sc=L.B, pr=0: A.cache_key(m, ph, sc, pr, {L.TRIAG: 4})
    assert k(m1) != k(m2) and k(m1) != k(m3)
    assert k(m1) != k(m1, pr=1) and k(m1) != k(m1, sc=L.N) and k(m1) != k(m1, ph=L.ADVECTION2D)
    assert k(m1) == A.cache_key(rectangle(8, 8, split=2, order=2, jitter=0.1, seed=3), L.BURGERS2D, L.B, 0, 4)


def test_cache_round_trip(tmp_path):
    p = str(tmp_path / "sub" / "cells.json")
    assert A.load_cache(p) == {}
    A.store_cache(p, "a", {"cells": 128})
    A.store_cache(p, "b", {"cells": 64})
    assert A.load_cache(p) == {"a": {"cells": 128}, "b": {"cells": 64}}
    with open(p, "w") as f:
        f.write("not json")
    assert A.load_cache(p) == {}
