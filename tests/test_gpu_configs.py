"""The benchmark configurations at their full sizes, as parity cases (the
uneven batch has its own module, test_gpu_c4.py).

* C1 (64 circuits, 10 qubits, 600 symbols): every expectation against the
  complex128 oracle (1e-5).
* C2 (30 qubits, 5x6 grid, 14 cycles): beyond the oracle's reach, so checked
  through size-independent properties: the 1e6-shot sampled expectation of the
  49-term sum ZZ and the ZZ estimate from 1e6 samples both agree with the exact
  expectation within their statistical error; every sample is a valid 30-bit
  string.
* C3 (26 qubits, 416 symbols, all 8 parameter sets): the adjoint gradient
  against parameter shift on 32 columns (rotation gates: [E(t + pi/2) -
  E(t - pi/2)] / 2, exact for ry/rz) at 5e-5 and central differences
  (h = 1e-3) computed with the exact expectation op; TFQ's finite-difference
  gradient gates agree with the exact derivative.
"""

import math
import sys

import numpy as np
import pytest

from oracle import tfq_oracle as O
from quantum_amd import serialize as S
from quantum_amd import workloads as W

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-case] {msg}", file=sys.stderr, flush=True)


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def test_config1_full_batch_matches_oracle():
    from quantum_amd import tfq_simulate_ops as sim
    circuits, names, values, ops = W.config1_line_rotations()
    progs = S.convert_to_tensor(circuits)
    ps = S.convert_to_tensor(ops)
    _say("C1 64 x 10 qubits")
    got = sim.tfq_simulate_expectation(progs, names, values, ps).cpu().numpy()
    want = O.simulate_expectation(list(progs), names, values, ps)
    np.testing.assert_allclose(got, want, atol=1e-5)


@pytest.fixture(scope="module")
def grid30():
    circuit, zz = W.config2_grid()
    return S.convert_to_tensor([circuit]), S.convert_to_tensor([[zz]]), len(zz.terms)


def test_config2_sampled_expectation_matches_exact(grid30):
    from quantum_amd import tfq_simulate_ops as sim
    progs, ps, nterms = grid30
    _say("C2 exact <sum ZZ> at 30 qubits")
    exact = float(sim.tfq_simulate_expectation(progs, [], np.zeros((1, 0)), ps).cpu()[0, 0])
    _say("C2 sampled expectation, 1e6 shots per term")
    est = float(sim.tfq_simulate_sampled_expectation(progs, [], np.zeros((1, 0)), ps,
                                                      np.full((1, 1), 10 ** 6, np.int32),
                                                      seed=1234).cpu()[0, 0])
    # Independent draws per term: std of the sum <= sqrt(49) / sqrt(1e6) = 7e-3.
    assert abs(est - exact) < 5 * math.sqrt(nterms) / 1e3, (est, exact)


def test_config2_samples_reproduce_the_zz_expectation(grid30):
    import torch
    from quantum_amd import tfq_simulate_ops as sim
    from quantum_amd import workloads as Wl
    progs, ps, nterms = grid30
    exact = float(sim.tfq_simulate_expectation(progs, [], np.zeros((1, 0)), ps).cpu()[0, 0])
    _say("C2 1e6 samples of 30 qubits")
    s = sim.tfq_simulate_samples(progs, [], np.zeros((1, 0)), [10 ** 6], seed=99)
    assert tuple(s.shape) == (1, 10 ** 6, 30)
    bits = s[0]
    assert bool(((bits == 0) | (bits == 1)).all())
    # Qubits are sorted (row, col); sample column j is qubit j.
    z = 1.0 - 2.0 * bits.to(torch.float32)
    circuit, zz = Wl.config2_grid()
    qs = sorted({q for term in zz.terms for q in term.paulis}, key=lambda q: (q.row, q.col))
    assert len(qs) == 30
    col = {q: j for j, q in enumerate(qs)}
    est = 0.0
    for term in zz.terms:
        a, b = [col[q] for q in term.paulis]
        est += float(term.coefficient.real) * float((z[:, a] * z[:, b]).mean())
    # Terms share shots; |sum| error <= sum of per-term errors (each std <= 1e-3).
    assert abs(est - exact) < 5 * nterms / 1e3, (est, exact)


@pytest.fixture(scope="module")
def hea26():
    circuit, names, values, ham = W.config3_hea()
    B = values.shape[0]
    return circuit, S.convert_to_tensor([circuit] * B), names, values, ham


def test_config3_adjoint_matches_parameter_shift_and_central_difference(hea26):
    """All 8 parameter sets; 32 symbol columns spread over the 8 layers (ry of
    qubits 0, 13 and 25, rz of qubit 7, in every layer).  For ry/rz the two-term
    parameter shift [E(t + pi/2) - E(t - pi/2)] / 2 is the exact derivative,
    so the analytic adjoint must match it to the float32 resolution of two
    26-qubit expectations of a 51-term observable (|E| <= 43.2): 5e-5.  All
    shifted sets run as one 512-row batch."""
    from quantum_amd import tfq_adj_grad_op as adj
    from quantum_amd import tfq_simulate_ops as sim
    circuit, progs, names, values, ham = hea26
    B = values.shape[0]
    ps = S.convert_to_tensor([[ham]] * B)
    ones = np.ones((B, 1), np.float32)
    _say("C3 adjoint gradient, 26 qubits, 8 x 416 symbols")
    g = adj.tfq_adj_grad(progs, names, values, ps, ones, derivative="analytic").cpu().numpy()
    g_tfq = adj.tfq_adj_grad(progs, names, values, ps, ones).cpu().numpy()
    assert g.shape == (B, 416)
    # TFQ's float32 finite-difference gradient gates track the exact derivative.
    np.testing.assert_allclose(g_tfq, g, atol=5e-3)
    cols = sorted({l * 52 + o for l in range(8) for o in (0, 13, 25, 26 + 7)})
    assert len(cols) == 32
    shifted, where = [], []
    for c in cols:
        for sign in (1, -1):
            v = values.astype(np.float64).copy()
            v[:, c] += sign * math.pi / 2
            shifted.append(v.astype(np.float32))
            where.append((c, sign))
    vals = np.concatenate(shifted)
    R = vals.shape[0]
    _say(f"C3 parameter shift: {R} shifted 26-qubit expectations in one call")
    e = sim.tfq_simulate_expectation(S.convert_to_tensor([circuit] * R), names, vals,
                                     S.convert_to_tensor([[ham]] * R)).cpu().numpy()[:, 0]
    e = e.reshape(len(cols), 2, B)
    shift = (e[:, 0, :] - e[:, 1, :]).T / 2
    err = float(np.abs(g[:, cols] - shift).max())
    _say(f"C3 max |adjoint - parameter shift| = {err:.2e}")
    np.testing.assert_allclose(g[:, cols], shift, atol=5e-5)
    for c in cols[:2]:
        h = 1e-3
        vp = values.astype(np.float64).copy()
        vm = vp.copy()
        vp[:, c] += h
        vm[:, c] -= h
        _say(f"C3 central difference, symbol {c}")
        ep = sim.tfq_simulate_expectation(progs, names, vp.astype(np.float32), ps).cpu().numpy()
        em = sim.tfq_simulate_expectation(progs, names, vm.astype(np.float32), ps).cpu().numpy()
        # float32 expectations: |dE| resolution ~1e-6 / (2h) plus the symbol's
        # float32 rounding.
        np.testing.assert_allclose(g[:, c], (ep - em)[:, 0] / (2 * h), atol=2e-2)
