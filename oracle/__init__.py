"""CPU oracle for the circuit-simulation op family.

TEST INFRASTRUCTURE ONLY: only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import, call, link or execute anything
in this directory, and only as the checker.  The product (``quantum_amd``)
never does, and fails loudly when its HIP library is missing.

Contents:
* ``tfq_oracle.py``  complex128 numpy restatement of TFQ's six ops and the
                     inner product's gradient op;
* ``wire.py``        an independent proto3 decoder for Program / PauliSum;
* ``philox.py``      host replica of the device sampler's Philox4x32-10 stream;
* ``mlp.py``         float64 restatement of the VQE energy surrogate network;
* ``cpu_tfq.cc``     float32 C++ restatement of TFQ's CPU (qsim) schedule with
  ``cpu_tfq.py``     std::threads, timed as the CPU baseline by ``bench.py``
                     (built by ``__graft_entry__.build()`` into libcpu_tfq.so).

Parity status of each function (how it is pinned) is recorded in
``tests/test_oracle.py``; no function is left "parity unpinned" except the
fusion decisions of qsim's BasicGateFuser, which no reference test fixes.
"""
