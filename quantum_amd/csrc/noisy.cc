// Host planner of the noisy trajectory kernel (see noisy.h).
#include "noisy.h"

#include <string.h>

#include <algorithm>
#include <cmath>
#include <numeric>

namespace tfqx {
namespace {

size_t Align(size_t x) { return (x + 255) & ~size_t(255); }

// A channel as a kOpChannel op (layout in tfqx_internal.h).
DevOp ChannelOp(const Channel& ch, int ordinal, std::vector<cfloat>* mats) {
  DevOp d{};
  d.type = kOpChannel;
  d.k = 1;
  d.q[0] = uint8_t(ch.bit);
  d.nderiv = 1;
  d.cval = uint32_t(ch.nk) | (ch.mixture ? kChanMixture | kChanFirstIdentity : 0u);
  d.col[0] = ordinal;
  d.mat = uint32_t(mats->size());
  for (int k = 0; k < ch.nk; ++k)
    for (int e = 0; e < 4; ++e)
      mats->push_back(cfloat(float(ch.K[k][e].real()), float(ch.K[k][e].imag())));
  d.dmat[0] = uint32_t(mats->size());
  if (ch.mixture) {
    // Cumulative probabilities summed left to right in double, then float32.
    double cum = 0.0;
    for (int k = 0; k < ch.nk; ++k) {
      cum += ch.probs[k];
      mats->push_back(cfloat(float(cum), 0.0f));
    }
  } else {
    // E_k = K_k^dag K_k.
    for (int k = 0; k < ch.nk; ++k) {
      const cdouble* K = ch.K[k];
      for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 2; ++c) {
          const cdouble e = std::conj(K[0 * 2 + r]) * K[0 * 2 + c] + std::conj(K[1 * 2 + r]) * K[1 * 2 + c];
          mats->push_back(cfloat(float(e.real()), float(e.imag())));
        }
    }
  }
  return d;
}

}  // namespace

Status BuildNoisyPlan(int mode, const std::vector<Circuit>& circuits,
                      const std::vector<std::vector<std::vector<Pauli>>>& sums,
                      const int32_t* num_samples, int nshots, int64_t B, int64_t O,
                      std::unique_ptr<NoisyPlan>* out) {
  auto p = std::make_unique<NoisyPlan>();
  p->mode = mode;
  p->B = B;
  p->O = O;
  p->nshots = nshots;
  p->ident.assign(size_t(B * O), 0.0f);
  p->num_samples.assign(size_t(B * O), 0);
  if (num_samples) p->num_samples.assign(num_samples, num_samples + B * O);
  p->row_empty.assign(size_t(B), 0);
  FuseOptions fo;
  fo.max_fuse = 2;  // the register engine applies one- and two-qubit blocks
  std::vector<DevChunk> wg_chunks, wave_chunks;
  std::vector<double> cost;
  for (int64_t i = 0; i < B; ++i) {
    const Circuit& c = circuits[size_t(i)];
    p->max_nq = std::max(p->max_nq, c.num_qubits);
    // Identity terms (and rows of empty programs) are finished on the host side.
    if (!sums.empty() && !sums[size_t(i)].empty())
      for (int64_t j = 0; j < O; ++j)
        for (const auto& t : sums[size_t(i)][size_t(j)])
          if (t.ops.empty()) p->ident[size_t(i * O + j)] += t.coef;
    // (#679) programs without operations give -2 (tfq_noisy_expectation.cc:190-196).
    if (c.gates.empty() && c.channels.empty()) {
      p->row_empty[size_t(i)] = 1;
      continue;
    }
    const int n = c.num_qubits;
    if (n > kNoisyMaxQubits)
      return Status{12, "Noisy circuits of at most " + std::to_string(kNoisyMaxQubits) +
                            " qubits are supported, got " + std::to_string(n) + "."};
    const bool wave = n <= kWaveMaxQubits;
    DevCircuit dc{};
    dc.n = wave ? std::max(n, kWaveRegBits) : n;
    dc.row = int(i);
    dc.flags = wave ? kFlagWave : 0;
    dc.fwd_begin = int(p->ops.size());
    // Fused gate segments between channels, each channel an op of its own.
    int g0 = 0;
    for (size_t k = 0; k <= c.channels.size(); ++k) {
      const int g1 = k < c.channels.size() ? c.channels[k].gate_pos : int(c.gates.size());
      if (g1 > g0) {
        std::vector<FusedOp> fused;
        FuseGates(c, g0, g1, fo, &fused);
        for (const auto& f : fused) p->ops.push_back(FusedToDevOp(f, &p->mats));
      }
      if (k < c.channels.size()) p->ops.push_back(ChannelOp(c.channels[k], int(k), &p->mats));
      g0 = g1;
    }
    if (wave) InsertRemapsR(&p->ops, size_t(dc.fwd_begin), dc.n, kWaveRegBits);
    dc.fwd_end = int(p->ops.size());
    // Non-identity terms, slot by slot.  Sampled expectation: unit weight
    // (times i^#Y) with the coefficient's float bits in `pad`.
    dc.term_begin = int(p->terms.size());
    if (mode != kModeNoisySamples) {
      for (int64_t j = 0; j < O; ++j) {
        for (const auto& t : sums[size_t(i)][size_t(j)]) {
          if (t.ops.empty()) continue;
          if (mode == kModeNoisyExpectation) {
            p->terms.push_back(PauliToTerm(t, t.coef, int(j)));
          } else {
            DevTerm d = PauliToTerm(t, 1.0f, int(j));
            memcpy(&d.pad, &t.coef, sizeof(float));
            p->terms.push_back(d);
          }
        }
      }
    }
    dc.term_end = int(p->terms.size());
    const int ci = int(p->circuits.size());
    p->circuits.push_back(dc);
    int T = 0;
    if (mode == kModeNoisySamples) {
      T = nshots;
    } else {
      for (int64_t j = 0; j < O; ++j) T = std::max(T, p->num_samples[size_t(i * O + j)]);
    }
    p->trajectories += T;
    const double amps = std::ldexp(1.0, dc.n);
    cost.push_back(amps * double(dc.fwd_end - dc.fwd_begin + 1));
    if (wave) {
      const int G = 1 << (kWaveMaxQubits - dc.n);
      for (int t = 0; t < T; t += G) wave_chunks.push_back(DevChunk{ci, t, std::min(T, t + G), n});
    } else {
      for (int t = 0; t < T; ++t) wg_chunks.push_back(DevChunk{ci, t, t + 1, n});
    }
  }
  // Longest circuits first within each phase.
  auto by_cost = [&](const DevChunk& a, const DevChunk& b) {
    return cost[size_t(a.circuit)] > cost[size_t(b.circuit)];
  };
  std::stable_sort(wg_chunks.begin(), wg_chunks.end(), by_cost);
  std::stable_sort(wave_chunks.begin(), wave_chunks.end(), by_cost);
  p->chunks = wg_chunks;
  p->nblock = int(wg_chunks.size());
  p->chunks.insert(p->chunks.end(), wave_chunks.begin(), wave_chunks.end());
  NoisyLayout(p.get());
  *out = std::move(p);
  return Status();
}

void NoisyLayout(NoisyPlan* p) {
  size_t off = 0;
  p->off_counters = off; off = Align(off + 256);
  p->off_circuits = off; off = Align(off + p->circuits.size() * sizeof(DevCircuit));
  p->off_ops = off; off = Align(off + p->ops.size() * sizeof(DevOp));
  p->off_mats = off; off = Align(off + (p->mats.size() + kMatCache) * sizeof(cfloat));
  p->off_terms = off; off = Align(off + p->terms.size() * sizeof(DevTerm));
  p->off_chunks = off; off = Align(off + p->chunks.size() * sizeof(DevChunk));
  p->off_ident = off; off = Align(off + p->ident.size() * sizeof(float));
  p->off_nsamp = off; off = Align(off + p->num_samples.size() * sizeof(int32_t));
  p->off_empty = off; off = Align(off + p->row_empty.size() * sizeof(int32_t));
  // The accumulators start at zero (the upload copies zeros) and every launch
  // leaves them so.
  p->off_nacc = off; off = Align(off + std::max<size_t>(1, p->ident.size()) * sizeof(double));
  p->image = off;
  p->grid = std::max(1, std::min(kNoisyGrid, int(p->chunks.size())));
  int lds_nq = kWaveMaxQubits + 3;
  for (int c = 0; c < p->nblock; ++c)
    if (p->chunks[size_t(c)].nq <= kNoisyLdsMaxQubits) lds_nq = std::max(lds_nq, p->chunks[size_t(c)].nq);
  p->lds_log2 = lds_nq;
  if (p->max_nq > kNoisyLdsMaxQubits) {
    p->state_log2 = p->max_nq;
    p->off_states = off;
    off = Align(off + size_t(p->grid) * (size_t(1) << p->state_log2) * sizeof(cfloat));
  }
  p->total = off;
  p->staging.assign(p->image, 0);
  auto put = [&](size_t o, const void* src, size_t n) {
    if (n) memcpy(p->staging.data() + o, src, n);
  };
  put(p->off_circuits, p->circuits.data(), p->circuits.size() * sizeof(DevCircuit));
  put(p->off_ops, p->ops.data(), p->ops.size() * sizeof(DevOp));
  put(p->off_mats, p->mats.data(), p->mats.size() * sizeof(cfloat));
  put(p->off_terms, p->terms.data(), p->terms.size() * sizeof(DevTerm));
  put(p->off_chunks, p->chunks.data(), p->chunks.size() * sizeof(DevChunk));
  put(p->off_ident, p->ident.data(), p->ident.size() * sizeof(float));
  put(p->off_nsamp, p->num_samples.data(), p->num_samples.size() * sizeof(int32_t));
  put(p->off_empty, p->row_empty.data(), p->row_empty.size() * sizeof(int32_t));
}

}  // namespace tfqx
