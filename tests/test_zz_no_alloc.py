"""tfqx_plan_launch_noalloc: the allocation-free launch (static device scratch)
against the workspace launch of the same plan, through the C-ABI, on two
calls in a row, the second with a smaller batch; plus its size limit.

Expectations: bit-identical (same kernels, same reduction order).  Adjoint
gradients: FP64 atomics whose order varies between launches, so 1e-6."""

import sys

import numpy as np
import pytest

from quantum_amd import _lib
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


def _plans(batch, seed):
    from quantum_amd import tfq_adj_grad_op as A
    from quantum_amd.tfq_simulate_ops import _common, _plan_args
    circuits, names, values, ops = W.config1_line_rotations(seed=seed, batch=batch)
    progs, pn, pv = _common(S.convert_to_tensor(circuits), names, values)
    ps = S.convert_to_tensor(ops)
    P, N, args = _plan_args(progs, pn, pv)
    Ps = _lib.StringTensor(ps)
    exp = _lib.Plan(_lib.lib().tfqx_plan_simulate_expectation, *args, Ps.ptr, Ps.off, ps.shape[1])
    g = np.random.default_rng(seed).uniform(-1, 1, (batch, ps.shape[1])).astype(np.float32)
    adj, shape = A.plan_adj_grad(progs, pn, pv, ps, g, with_expectation=True)
    return (exp, (batch, ps.shape[1])), (adj, shape)


def _both(plan, shape):
    import torch
    dev = torch.device("cuda", torch.cuda.current_device())
    stream = torch.cuda.current_stream(dev).cuda_stream
    ref = torch.empty(shape, dtype=torch.float32, device=dev)
    plan.upload(dev, stream)
    plan.launch(ref, stream)
    got = torch.full(shape, float("nan"), dtype=torch.float32, device=dev)
    plan.launch_noalloc(got, stream)
    torch.cuda.synchronize()
    return got.cpu().numpy(), ref.cpu().numpy()


def test_noalloc_launch_equals_workspace_launch_twice_then_smaller():
    for batch, seed in ((64, 1000), (9, 77)):
        (exp, es), (adj, ash) = _plans(batch, seed)
        _say(f"noalloc expectation batch={batch}")
        got, ref = _both(exp, es)
        np.testing.assert_array_equal(got, ref)
        _say(f"noalloc expectation + adjoint batch={batch}")
        got, ref = _both(adj, ash)
        B, O = es
        np.testing.assert_array_equal(got[:B * O], ref[:B * O])
        np.testing.assert_allclose(got[B * O:], ref[B * O:], atol=1e-6)


def test_noalloc_rejects_plans_over_its_limit():
    import torch
    from quantum_amd import tfq_adj_grad_op as A
    circuits, names, values, psums, weights = W.config4_uneven(count=64)
    progs = S.convert_to_tensor(circuits)
    ps = S.convert_to_tensor(psums)
    plan, shape = A.plan_adj_grad(progs, names, values, ps, weights, with_expectation=True)
    assert plan.workspace_bytes > 32 * 2 ** 20
    out = torch.empty(shape, dtype=torch.float32, device="cuda")
    with pytest.raises(_lib.InvalidArgumentError, match="static scratch"):
        plan.launch_noalloc(out, torch.cuda.current_stream().cuda_stream)
