"""The simulation ops, with the call signatures of TFQ's
``core/ops/tfq_simulate_ops.py:22-135``.

Each function takes the same inputs as its TFQ counterpart -- serialized
programs (``convert_to_tensor`` output, or ``Circuit`` objects), symbol names,
symbol values (cast to float32, ``tfq_simulate_ops.py:44-45``) and serialized
Pauli sums -- and returns a ``torch`` tensor on the current HIP device with the
reference's dtype, shape and sentinel values.  The work runs in ``libtfqx.so``
(parse/fuse on the host, kernels on the GPU); there is no CPU fallback.
"""

import numpy as np

from . import _lib
from ._lib import InvalidArgumentError, Plan, StringTensor, np_ptr
from . import serialize


def _torch():
    import torch
    return torch


def as_string_tensor(x, rank, name):
    """Object ndarray of bytes from serialized strings or circuit/Pauli objects."""
    if isinstance(x, np.ndarray) and x.dtype == object:
        arr = x
    elif isinstance(x, np.ndarray) and x.dtype.kind in "SU":
        arr = np.array([[v for v in r] for r in x], dtype=object) if x.ndim == 2 else \
            np.array(list(x), dtype=object)
        arr = arr.reshape(x.shape)
    else:
        arr = serialize.convert_to_tensor(list(x)) if not isinstance(x, (bytes, str)) else \
            np.array(x, dtype=object)
        if arr.dtype != object:
            arr = arr.astype(object)
        if isinstance(x, (list, tuple)) and len(x) == 0:
            arr = np.empty((0,) * rank, dtype=object)
    if arr.ndim != rank:
        raise InvalidArgumentError(f"{name} must be rank {rank}. Got rank {arr.ndim}.")
    return arr


def as_values(v, name="symbol_values", rank=2, dtype=np.float32):
    torch = _torch()
    if isinstance(v, torch.Tensor):
        v = v.detach().to("cpu").numpy()
    arr = np.asarray(v)
    if arr.dtype.kind in "SUO" and any(isinstance(x, (str, bytes)) for x in arr.reshape(-1)):
        # tf.cast of a string tensor (UnimplementedError in the reference).
        raise NotImplementedError("Cast string to float is not supported")
    if arr.dtype == object:
        arr = arr.astype(np.float64)
    if np.iscomplexobj(arr):
        arr = arr.real
    arr = np.ascontiguousarray(arr, dtype=dtype)
    if arr.ndim != rank:
        raise InvalidArgumentError(f"{name} must be rank {rank}. Got rank {arr.ndim}.")
    return arr


def as_names(names):
    if isinstance(names, np.ndarray):
        names = list(names.reshape(-1)) if names.ndim <= 1 else None
        if names is None:
            raise InvalidArgumentError("symbol_names must be rank 1.")
    out = []
    for n in list(names):
        if isinstance(n, bytes):
            out.append(n.decode("utf-8"))
        elif isinstance(n, str):
            out.append(n)
        else:
            # A string tensor cannot hold numbers (TypeError from
            # tf.convert_to_tensor in the reference).
            raise TypeError(f"Cannot convert {list(names)!r} to EagerTensor of dtype string")
    return np.array(out, dtype=object)


def _device_and_stream():
    torch = _torch()
    if not torch.cuda.is_available():
        raise RuntimeError("The tfqx ops need a HIP device (torch.cuda.is_available() is False).")
    dev = torch.device("cuda", torch.cuda.current_device())
    return dev, torch.cuda.current_stream(dev).cuda_stream


def _check_batch(programs, values):
    if programs.shape[0] != values.shape[0]:
        raise InvalidArgumentError(
            f"Number of circuits and symbol_values do not match. Got {programs.shape[0]} circuits"
            f" and {values.shape[0]} symbol values.")


def _check_names(names, values):
    if values.shape[1] != len(names):
        raise InvalidArgumentError("Input symbol names and value sizes do not match.")


def _common(programs, symbol_names, symbol_values):
    programs = as_string_tensor(programs, 1, "programs")
    names = as_names(symbol_names)
    values = as_values(symbol_values)
    if values.shape[0] == 0 and programs.shape[0] == 0 and values.shape[1] != len(names):
        values = np.zeros((0, len(names)), np.float32)
    _check_batch(programs, values)
    _check_names(names, values)
    return programs, names, values


def _run(plan, out_shape, dtype):
    torch = _torch()
    dev, stream = _device_and_stream()
    out = torch.empty(out_shape, dtype=dtype, device=dev)
    if out.numel() == 0 and plan.workspace_bytes <= 256:
        return out
    plan.upload(dev, stream)
    plan.launch(out, stream)
    return out


def _plan_args(programs, names, values):
    P = StringTensor(programs)
    N = StringTensor(names)
    return P, N, [P.ptr, P.off, P.n, N.ptr, N.off, N.n, np_ptr(values)]


def plan_expectation(programs, symbol_names, symbol_values, pauli_sums):
    programs, names, values = _common(programs, symbol_names, symbol_values)
    pauli_sums = as_string_tensor(pauli_sums, 2, "pauli_sums")
    if pauli_sums.shape[0] != programs.shape[0]:
        raise InvalidArgumentError(
            f"Number of circuits and PauliSums do not match. Got {programs.shape[0]} circuits "
            f"and {pauli_sums.shape[0]} paulisums.")
    P, N, args = _plan_args(programs, names, values)
    S = StringTensor(pauli_sums)
    plan = Plan(_lib.lib().tfqx_plan_simulate_expectation, *args, S.ptr, S.off,
                pauli_sums.shape[1])
    return plan, (programs.shape[0], pauli_sums.shape[1])


def tfq_simulate_expectation(programs, symbol_names, symbol_values, pauli_sums):
    """Expectation values ``[batch, n_ops]`` (float32); -2 for empty programs.

    A call with the same programs, names and Pauli sums as a recent one reuses
    its plan (``plan_cache``): the new symbol values are uploaded and evaluated
    into the plan's matrices on the device."""
    from . import plan_cache
    programs, names, values = _common(programs, symbol_names, symbol_values)
    pauli_sums = as_string_tensor(pauli_sums, 2, "pauli_sums")
    _, stream = _device_and_stream()
    key = ("expectation", plan_cache.digest(programs, pauli_sums), tuple(names),
           programs.shape, pauli_sums.shape, stream)
    shape = (programs.shape[0], pauli_sums.shape[1])

    def build():
        return plan_expectation(programs, names, values, pauli_sums)[0]

    return plan_cache.run(key, build, lambda plan: plan.stage_values(values),
                          lambda plan: _run(plan, shape, _torch().float32))


def tfq_simulate_state(programs, symbol_names, symbol_values):
    """Final states ``[batch, 2**max_nq]`` (complex64), padded with -2."""
    programs, names, values = _common(programs, symbol_names, symbol_values)
    P, N, args = _plan_args(programs, names, values)
    plan = Plan(_lib.lib().tfqx_plan_simulate_state, *args)
    return _run(plan, (programs.shape[0], 2 ** plan.max_num_qubits), _torch().complex64)


def _num_samples_scalar(num_samples):
    ns = np.asarray(_torch().as_tensor(num_samples).cpu().numpy() if hasattr(num_samples, "cpu")
                    else num_samples)
    if ns.ndim != 1:
        raise InvalidArgumentError(f"num_samples must be rank 1. Got rank {ns.ndim}.")
    if ns.shape[0] != 1:
        raise InvalidArgumentError(f"num_samples must contain 1 element. Got {ns.shape[0]}.")
    return int(ns[0])


def _seed(seed):
    if seed is None:
        seed = int(np.random.SeedSequence().generate_state(1, dtype=np.uint64)[0])
    return int(seed) & 0xFFFFFFFFFFFFFFFF


def tfq_simulate_samples(programs, symbol_names, symbol_values, num_samples, seed=None):
    """Bitstrings ``[batch, num_samples, max_nq]`` (int8), leading -2 padding.

    ``seed`` fixes the Philox stream (the reference draws a fresh seed per call)."""
    programs, names, values = _common(programs, symbol_names, symbol_values)
    n = _num_samples_scalar(num_samples)
    P, N, args = _plan_args(programs, names, values)
    plan = Plan(_lib.lib().tfqx_plan_simulate_samples, *args, n, _seed(seed))
    return _run(plan, (programs.shape[0], n, plan.max_num_qubits), _torch().int8)


def tfq_simulate_sampled_expectation(programs, symbol_names, symbol_values, pauli_sums,
                                     num_samples, seed=None):
    """Sampled expectations ``[batch, n_ops]`` with ``num_samples[i][j]`` shots
    per term of ``pauli_sums[i][j]``."""
    programs, names, values = _common(programs, symbol_names, symbol_values)
    pauli_sums = as_string_tensor(pauli_sums, 2, "pauli_sums")
    ns = as_values(num_samples, "num_samples", 2, np.int32)
    if pauli_sums.shape[0] != programs.shape[0]:
        raise InvalidArgumentError(
            f"Number of circuits and PauliSums do not match. Got {programs.shape[0]} circuits "
            f"and {pauli_sums.shape[0]} paulisums.")
    if ns.shape != pauli_sums.shape:
        raise InvalidArgumentError(
            f"num_samples and pauli_sums do not match. Got {list(ns.shape)} and "
            f"{list(pauli_sums.shape)}.")
    P, N, args = _plan_args(programs, names, values)
    S = StringTensor(pauli_sums)
    plan = Plan(_lib.lib().tfqx_plan_simulate_sampled_expectation, *args, S.ptr, S.off,
                pauli_sums.shape[1], np_ptr(ns), _seed(seed))
    return _run(plan, pauli_sums.shape, _torch().float32)


def padded_to_ragged(masked_state):
    """``tfq_utility_ops.padded_to_ragged`` (``core/ops/tfq_utility_ops.py:69-85``):
    drop entries whose magnitude is >= 1.1 (the -2 padding)."""
    torch = _torch()
    x = masked_state.detach().cpu()
    mag = x.abs() if x.is_complex() else x.abs().float()
    rows = []
    for i in range(x.shape[0]):
        keep = mag[i] < 1.1
        rows.append(x[i][keep] if x.dim() == 2 else x[i][keep.all(dim=-1)])
    return rows


def padded_to_ragged2d(masked_state):
    """Samples: drop padded bit columns per row (``tfq_utility_ops.py:88-103``)."""
    out = []
    x = masked_state.detach().cpu()
    for i in range(x.shape[0]):
        cols = (x[i] != -2).all(dim=0) if x.shape[1] else x.new_zeros(x.shape[2], dtype=bool)
        out.append(x[i][:, cols])
    return out
