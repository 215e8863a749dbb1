"""Host logic of the op-path plan cache: digests of string tensors, the
identity memo, LRU eviction and settings generations (no device needed)."""

import numpy as np

from quantum_amd import plan_cache


def _arr(items):
    a = np.empty(len(items), dtype=object)
    a[:] = items
    return a


def test_digest_depends_on_contents_not_objects():
    plan_cache.clear()
    a = _arr([b"abc", b"de", b""])
    b = _arr([bytes(bytearray(b"abc")), bytes(bytearray(b"de")), b""])
    assert plan_cache.digest(a) == plan_cache.digest(b)
    assert plan_cache.digest(a) == plan_cache.digest(a)  # memo hit
    # Same concatenation, different boundaries.
    assert plan_cache.digest(_arr([b"ab", b"cde", b""])) != plan_cache.digest(a)
    c = a.copy()
    c[1] = b"dx"
    assert plan_cache.digest(c) != plan_cache.digest(a)
    assert plan_cache.digest(a, None) != plan_cache.digest(a)
    assert plan_cache.digest(a.reshape(3, 1)) == plan_cache.digest(a)


def test_replacing_an_element_in_place_changes_the_digest():
    plan_cache.clear()
    a = _arr([b"x" * 100, b"y" * 100])
    d0 = plan_cache.digest(a)
    a[0] = b"z" * 100
    assert plan_cache.digest(a) != d0


def test_string_elements_are_hashed_without_the_memo():
    plan_cache.clear()
    a = _arr(["abc", "de"])
    assert plan_cache.digest(a) == plan_cache.digest(_arr(["abc", "de"]))
    assert len(plan_cache._MEMO) == 0


class _FakePlan:
    def __init__(self, streaming=0):
        self.streaming = streaming
        self.rebinds = 0

    def stats(self):
        return {"streaming": self.streaming}


def test_run_builds_once_then_rebinds_and_evicts_lru():
    plan_cache.clear()
    plan_cache.set_capacity(2)
    built = []

    def build():
        built.append(_FakePlan())
        return built[-1]

    def rebind(p):
        p.rebinds += 1

    try:
        for key in ["a", "a", "b", "a", "c", "b"]:
            plan_cache.run((key,), build, rebind, lambda p: p)
        # a built, a rebound, b built, a rebound, c built (evicts b), b built.
        assert len(built) == 4
        assert built[0].rebinds == 2
        assert len(plan_cache._CACHE) == 2
        plan_cache.bump_generation()
        assert len(plan_cache._CACHE) == 0
        plan_cache.run(("a",), build, rebind, lambda p: p)
        assert len(built) == 5
    finally:
        plan_cache.set_capacity(8)
        plan_cache.clear()


def test_streaming_plans_are_not_cached_and_capacity_zero_disables():
    plan_cache.clear()
    n = [0]

    def build(streaming):
        n[0] += 1
        return _FakePlan(streaming)

    plan_cache.run(("s",), lambda: build(1), lambda p: None, lambda p: p)
    plan_cache.run(("s",), lambda: build(1), lambda p: None, lambda p: p)
    assert n[0] == 2 and len(plan_cache._CACHE) == 0
    plan_cache.set_capacity(0)
    try:
        plan_cache.run(("t",), lambda: build(0), lambda p: None, lambda p: p)
        plan_cache.run(("t",), lambda: build(0), lambda p: None, lambda p: p)
        assert n[0] == 4 and len(plan_cache._CACHE) == 0
    finally:
        plan_cache.set_capacity(8)
