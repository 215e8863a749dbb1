// Launch interface of the persistent LDS-resident circuit kernel.
#pragma once

#include <hip/hip_runtime.h>

#include "tfqx_internal.h"

namespace tfqx {

struct BatchArgs {
  int mode;            // BatchMode
  int kmax;            // widest fused op in the program (selects the kernel variant)
  const DevCircuit* circuits;
  int ncirc;
  int nblock;          // circuits [0, nblock) run per workgroup, the rest per wavefront
  int nslots;          // O: expectation columns (K for inner products)
  int nsym;            // S: gradient columns
  int lds_log2;        // LDS holds 2^lds_log2 amplitudes per state buffer
  int lds_amps;
  const DevOp* ops;
  const float2* mats;
  const DevTerm* terms;
  const DevOther* others;
  const DevPass* passes;     // tiled circuits
  float2* scratch;           // tiled circuits: per-workgroup psi | lambda
  long long scratch_stride;  // float2 elements per workgroup slice
  const float* ident;  // [B*O]
  float* exp_out;      // [B*O] expectation / sampled expectation
  float* grad_out;     // [B*S] adjoint
  double* grad_acc;    // [B*S] FP64 scratch (adjoint)
  float2* state_out;   // [B*stride] state
  long long state_stride;
  signed char* samples_out;  // [B*nshots*max_nq]
  int nshots;
  int max_nq;
  const int* num_samples;    // [B*O] sampled expectation
  unsigned long long seed;
  float2* ip_out;      // [B*K] inner products
  float2 rot[8];       // Y^-0.5 (4 entries) then X^0.5 (4 entries): Z-basis rotations
  int* queue;          // work-queue head (zero between launches)
  int* done;           // workgroups finished (zero between launches)
  int* wqueue;         // wave-mode queue head (zero between launches)
};

// Largest circuit the LDS kernel holds (per mode): 2^14 amplitudes (128 KiB)
// with one state buffer; 2 x 2^13 with two (adjoint lambda, inner-product
// paired state, sampled-expectation rotated copy).
constexpr int kLdsMaxQubitsOneBuffer = 14;
constexpr int kLdsMaxQubitsTwoBuffers = 13;
// Wider circuits (up to kTiledMaxQubits) still run inside the batch kernel:
// their state lives in a per-workgroup global scratch slice and is streamed
// through LDS tile by tile.  The grid of such launches is capped so the
// scratch stays bounded.
constexpr int kTiledMaxQubits = 18;
constexpr int kTiledGridCap = 512;

int BatchBuffers(int mode);
size_t BatchLdsBytes(int lds_log2, int mode);
hipError_t LaunchBatch(const BatchArgs& args, int grid, hipStream_t stream);
int BatchMaxGrid(int lds_log2, int mode, int kmax);

}  // namespace tfqx
