"""Op getters' ``quantum_concurrent`` keyword (``circuit_execution_ops.py:41-44,
133-142`` of the reference): ``False`` serialises the getter's ops under the
global op lock, ``True`` leaves them free to overlap; non-bool values raise
``TypeError``; the default follows ``set_quantum_concurrent_op_mode``.  The
simulate ops are replaced by a recorder, so these run on the CPU."""

import threading
import time

import pytest

from quantum_amd import circuit_execution_ops as ceo


class _Recorder:
    def __init__(self):
        self.lock = threading.Lock()
        self.active = 0
        self.peak = 0

    def __call__(self, *args, **kwargs):
        with self.lock:
            self.active += 1
            self.peak = max(self.peak, self.active)
        time.sleep(0.05)
        with self.lock:
            self.active -= 1
        return "ok"


@pytest.fixture
def recorder(monkeypatch):
    rec = _Recorder()
    for name in ("tfq_simulate_expectation", "tfq_simulate_sampled_expectation",
                 "tfq_simulate_samples", "tfq_simulate_state"):
        monkeypatch.setattr(ceo.sim, name, rec)
    yield rec
    ceo.set_quantum_concurrent_op_mode(True)


_GETTERS = [
    (ceo.get_expectation_op, 4),
    (ceo.get_sampled_expectation_op, 5),
    (ceo.get_sampling_op, 4),
    (ceo.get_state_op, 3),
]


def _run_threads(ops_and_arity, threads=4):
    start = threading.Barrier(threads)
    results = []

    def body(op, k):
        start.wait()
        results.append(op(*([None] * k)))

    ts = [threading.Thread(target=body, args=ops_and_arity[i % len(ops_and_arity)])
          for i in range(threads)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    return results


@pytest.mark.parametrize("getter,arity", _GETTERS)
def test_non_concurrent_ops_never_overlap(recorder, getter, arity):
    op = getter(quantum_concurrent=False)
    assert _run_threads([(op, arity)]) == ["ok"] * 4
    assert recorder.peak == 1


@pytest.mark.parametrize("getter,arity", _GETTERS)
def test_concurrent_ops_overlap(recorder, getter, arity):
    op = getter(quantum_concurrent=True)
    assert _run_threads([(op, arity)]) == ["ok"] * 4
    assert recorder.peak > 1


def test_lock_is_shared_across_getters(recorder):
    ops = [(g(quantum_concurrent=False), k) for g, k in _GETTERS]
    _run_threads(ops, threads=8)
    assert recorder.peak == 1


def test_default_follows_global_mode(recorder):
    ceo.set_quantum_concurrent_op_mode(False)
    assert ceo.get_quantum_concurrent_op_mode() is False
    _run_threads([(ceo.get_expectation_op(), 4)])
    assert recorder.peak == 1
    ceo.set_quantum_concurrent_op_mode(True)
    _run_threads([(ceo.get_expectation_op(), 4)])
    assert recorder.peak > 1


@pytest.mark.parametrize("getter,_", _GETTERS)
def test_non_bool_is_rejected(getter, _):
    with pytest.raises(TypeError, match="quantum_concurrent must be type bool"):
        getter(quantum_concurrent="junk")
