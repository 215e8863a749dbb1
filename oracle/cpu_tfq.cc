// CPU restatement of TFQ's qsim path (test and measurement infrastructure
// only: the reported CPU baseline of bench.py; the product never links it).
//
// Follows the structure of TFQ's CPU ops, in float32 complex like qsim's
// state space:
//   expectation   simulate the (fused) gate list, then per Pauli term copy the
//                 state, apply the term's Paulis and take the real inner
//                 product (ComputeExpectationQsim, core/src/util_qsim.h:142-188);
//                 identity terms add their coefficient.
//   adjoint       simulate, lambda = sum_j g_j sum_t c_t P_t psi with
//                 |g_j c_t| < 1e-5 dropped (AccumulateOperators,
//                 util_qsim.h:363-414), then for gates in reverse:
//                 psi <- G^dag psi; per symbol: scratch = dG psi (control-
//                 satisfied amplitudes only), grad += 2 Re<scratch|lambda>;
//                 lambda <- G^dag lam (core/ops/tfq_adj_grad_op.cc:214-276).
// Two schedules, as TFQ's ops choose them: ComputeSmall runs circuits in
// parallel over std::threads, one circuit per worker (util_qsim.h:46-131);
// ComputeLarge runs circuits one after another with every thread on the one
// state (OpenMP loops, the analogue of qsim's parallel For), which TFQ picks
// when the widest circuit has >= 26 qubits (>= 25 for the adjoint) or the
// batch holds one circuit (tfq_simulate_expectation_op.cc:120-126,
// tfq_adj_grad_op.cc:161-169).  Samples follow qsim's sampler: FP64 norm,
// sorted uniforms times the norm, one linear scan of the cumulative mass.
//
// Gate conventions: bits[0] is the matrix-index MSB; matrices are row-major
// complex64.  The Python side (oracle/cpu_tfq.py) lowers programs with the
// oracle's parser and supplies matrices and derivative matrices.
#include <stdint.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <complex>
#include <cstring>
#include <thread>
#include <vector>

namespace {

using c64 = std::complex<float>;

struct Gate {
  int32_t nt, b0, b1, mat, nderiv, dmat0, dmat1, col0, col1;
  uint64_t cmask, cval;
};

inline c64 mulc(c64 a, c64 b) {
  return c64(a.real() * b.real() - a.imag() * b.imag(), a.real() * b.imag() + a.imag() * b.real());
}

inline uint64_t ins0(uint64_t x, int b) { return ((x >> b) << (b + 1)) | (x & ((uint64_t(1) << b) - 1)); }

void load_matrix(const c64* M, int d, bool dag, c64* out) {
  for (int r = 0; r < d; ++r)
    for (int c = 0; c < d; ++c) out[r * d + c] = dag ? std::conj(M[c * d + r]) : M[r * d + c];
}

// s <- U s on the gate's targets (controls: amplitudes with (x & cmask) == cval).
// zero_others: amplitudes failing the control are set to zero (projection).
void apply(c64* s, int n, const Gate& g, const c64* U, bool zero_others = false) {
  const uint64_t half = uint64_t(1) << (n - g.nt);
  if (g.nt == 1) {
    const uint64_t bit = uint64_t(1) << g.b0;
    for (uint64_t i = 0; i < half; ++i) {
      const uint64_t x0 = ins0(i, g.b0), x1 = x0 | bit;
      if ((x0 & g.cmask) != g.cval) {
        if (zero_others) s[x0] = s[x1] = c64(0, 0);
        continue;
      }
      const c64 a0 = s[x0], a1 = s[x1];
      s[x0] = mulc(U[0], a0) + mulc(U[1], a1);
      s[x1] = mulc(U[2], a0) + mulc(U[3], a1);
    }
    return;
  }
  const int lo = g.b0 < g.b1 ? g.b0 : g.b1, hi = g.b0 < g.b1 ? g.b1 : g.b0;
  const uint64_t m0 = uint64_t(1) << g.b0, m1 = uint64_t(1) << g.b1;
  for (uint64_t i = 0; i < half; ++i) {
    const uint64_t base = ins0(ins0(i, lo), hi);
    uint64_t x[4];
    for (int m = 0; m < 4; ++m) x[m] = base | ((m >> 1) ? m0 : 0) | ((m & 1) ? m1 : 0);
    if ((base & g.cmask) != g.cval) {
      if (zero_others)
        for (int m = 0; m < 4; ++m) s[x[m]] = c64(0, 0);
      continue;
    }
    c64 a[4];
    for (int m = 0; m < 4; ++m) a[m] = s[x[m]];
    for (int r = 0; r < 4; ++r) {
      c64 acc(0, 0);
      for (int c = 0; c < 4; ++c) acc += mulc(U[r * 4 + c], a[c]);
      s[x[r]] = acc;
    }
  }
}

// s <- P s for one Pauli string (types 0 X, 1 Y, 2 Z on bits).
void apply_paulis(c64* s, int n, const int32_t* paulis, int np) {
  static const c64 X[4] = {c64(0, 0), c64(1, 0), c64(1, 0), c64(0, 0)};
  static const c64 Y[4] = {c64(0, 0), c64(0, -1), c64(0, 1), c64(0, 0)};
  static const c64 Z[4] = {c64(1, 0), c64(0, 0), c64(0, 0), c64(-1, 0)};
  for (int k = 0; k < np; ++k) {
    Gate g{};
    g.nt = 1;
    g.b0 = paulis[2 * k];
    const int t = paulis[2 * k + 1];
    apply(s, n, g, t == 0 ? X : (t == 1 ? Y : Z));
  }
}

double real_inner(const c64* a, const c64* b, uint64_t N) {
  double acc = 0;
  for (uint64_t i = 0; i < N; ++i) acc += double(a[i].real()) * b[i].real() + double(a[i].imag()) * b[i].imag();
  return acc;
}


// ------------------------------------------------------------------ state-parallel (ComputeLarge)
void apply_par(c64* s, int n, const Gate& g, const c64* U, bool zero_others, int nt) {
  const int64_t half = int64_t(1) << (n - g.nt);
  if (g.nt == 1) {
    const uint64_t bit = uint64_t(1) << g.b0;
#pragma omp parallel for num_threads(nt) schedule(static)
    for (int64_t i = 0; i < half; ++i) {
      const uint64_t x0 = ins0(uint64_t(i), g.b0), x1 = x0 | bit;
      if ((x0 & g.cmask) != g.cval) {
        if (zero_others) s[x0] = s[x1] = c64(0, 0);
        continue;
      }
      const c64 a0 = s[x0], a1 = s[x1];
      s[x0] = mulc(U[0], a0) + mulc(U[1], a1);
      s[x1] = mulc(U[2], a0) + mulc(U[3], a1);
    }
    return;
  }
  const int lo = g.b0 < g.b1 ? g.b0 : g.b1, hi = g.b0 < g.b1 ? g.b1 : g.b0;
  const uint64_t m0 = uint64_t(1) << g.b0, m1 = uint64_t(1) << g.b1;
#pragma omp parallel for num_threads(nt) schedule(static)
  for (int64_t i = 0; i < half; ++i) {
    const uint64_t base = ins0(ins0(uint64_t(i), lo), hi);
    uint64_t x[4];
    for (int m = 0; m < 4; ++m) x[m] = base | ((m >> 1) ? m0 : 0) | ((m & 1) ? m1 : 0);
    if ((base & g.cmask) != g.cval) {
      if (zero_others)
        for (int m = 0; m < 4; ++m) s[x[m]] = c64(0, 0);
      continue;
    }
    c64 a[4];
    for (int m = 0; m < 4; ++m) a[m] = s[x[m]];
    for (int r = 0; r < 4; ++r) {
      c64 acc(0, 0);
      for (int c = 0; c < 4; ++c) acc += mulc(U[r * 4 + c], a[c]);
      s[x[r]] = acc;
    }
  }
}

void apply_paulis_par(c64* s, int n, const int32_t* paulis, int np, int nt) {
  static const c64 X[4] = {c64(0, 0), c64(1, 0), c64(1, 0), c64(0, 0)};
  static const c64 Y[4] = {c64(0, 0), c64(0, -1), c64(0, 1), c64(0, 0)};
  static const c64 Z[4] = {c64(1, 0), c64(0, 0), c64(0, 0), c64(-1, 0)};
  for (int k = 0; k < np; ++k) {
    Gate g{};
    g.nt = 1;
    g.b0 = paulis[2 * k];
    const int t = paulis[2 * k + 1];
    apply_par(s, n, g, t == 0 ? X : (t == 1 ? Y : Z), false, nt);
  }
}

double real_inner_par(const c64* a, const c64* b, int64_t N, int nt) {
  double acc = 0;
#pragma omp parallel for num_threads(nt) schedule(static) reduction(+ : acc)
  for (int64_t i = 0; i < N; ++i)
    acc += double(a[i].real()) * b[i].real() + double(a[i].imag()) * b[i].imag();
  return acc;
}

void copy_par(c64* dst, const c64* src, int64_t N, int nt) {
#pragma omp parallel for num_threads(nt) schedule(static)
  for (int64_t i = 0; i < N; ++i) dst[i] = src[i];
}

void zero_state_par(c64* s, int64_t N, int nt) {
#pragma omp parallel for num_threads(nt) schedule(static)
  for (int64_t i = 0; i < N; ++i) s[i] = c64(0, 0);
  s[0] = c64(1, 0);
}

void axpy_par(c64* y, float w, const c64* x, int64_t N, int nt) {
#pragma omp parallel for num_threads(nt) schedule(static)
  for (int64_t i = 0; i < N; ++i) y[i] += w * x[i];
}

double philox_u(unsigned long long seed, unsigned long long shot, uint32_t s_lo, uint32_t s_hi) {
  uint32_t c0 = uint32_t(shot), c1 = uint32_t(shot >> 32), c2 = s_lo, c3 = s_hi;
  uint32_t k0 = uint32_t(seed), k1 = uint32_t(seed >> 32);
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = uint64_t(0xD2511F53u) * c0, p1 = uint64_t(0xCD9E8D57u) * c2;
    const uint32_t n0 = uint32_t(p1 >> 32) ^ c1 ^ k0, n2 = uint32_t(p0 >> 32) ^ c3 ^ k1;
    c0 = n0; c1 = uint32_t(p1); c2 = n2; c3 = uint32_t(p0);
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  const unsigned long long bits = (uint64_t(c0) << 32) | c1;
  return double(bits >> 11) * (1.0 / 9007199254740992.0);
}


// qsim-style sampler: FP64 norm, sorted uniforms u_k * norm, one scan of the
// cumulative mass (parallel over blocks of the state).  Indices come back in
// ascending order, as qsim returns them.
void sample_par(const c64* s, int64_t N, int64_t count, uint64_t seed, uint32_t s_lo, uint32_t s_hi,
                int nt, uint64_t* out) {
  const int nb = std::max(1, nt);
  std::vector<double> bsum(size_t(nb), 0.0);
  const int64_t chunk = (N + nb - 1) / nb;
#pragma omp parallel for num_threads(nt) schedule(static)
  for (int b = 0; b < nb; ++b) {
    double acc = 0;
    const int64_t e = std::min(N, (b + 1) * chunk);
    for (int64_t i = b * chunk; i < e; ++i) acc += std::norm(s[i]);
    bsum[size_t(b)] = acc;
  }
  double norm = 0;
  std::vector<double> bstart(size_t(nb) + 1, 0.0);
  for (int b = 0; b < nb; ++b) bstart[size_t(b) + 1] = bstart[size_t(b)] + bsum[size_t(b)];
  norm = bstart[size_t(nb)];
  std::vector<double> rs(static_cast<size_t>(count));
  for (int64_t k = 0; k < count; ++k) rs[size_t(k)] = philox_u(seed, uint64_t(k), s_lo, s_hi) * norm;
  std::sort(rs.begin(), rs.end());
  // Draws falling in block b are [first[b], first[b+1]).
  std::vector<int64_t> first(size_t(nb) + 1, count);
  for (int b = 0; b < nb; ++b)
    first[size_t(b)] = int64_t(std::lower_bound(rs.begin(), rs.end(), bstart[size_t(b)]) - rs.begin());
  first[size_t(nb)] = count;
#pragma omp parallel for num_threads(nt) schedule(static)
  for (int b = 0; b < nb; ++b) {
    int64_t k = first[size_t(b)];
    const int64_t kend = first[size_t(b) + 1];
    double cum = bstart[size_t(b)];
    const int64_t e = std::min(N, (b + 1) * chunk);
    for (int64_t i = b * chunk; i < e && k < kend; ++i) {
      cum += std::norm(s[i]);
      while (k < kend && rs[size_t(k)] < cum) out[k++] = uint64_t(i);
    }
    for (; k < kend; ++k) out[k] = uint64_t(e - 1);
  }
}

}  // namespace

extern "C" {

// circ[B][8]: n, gate_begin, gate_end, fgate_begin, fgate_end, term_begin, term_end, row.
// gates_i[G][9]: nt, b0, b1, mat, nderiv, dmat0, dmat1, col0, col1; gates_m[G][2]: cmask, cval.
// terms_i[T][3]: slot, pauli_begin, pauli_end; terms_c[T]: coefficient; paulis[P][2]: bit, type.
// mode 0: expectation -> exp_out[B][O] (fused gates); mode 1: adjoint -> grad_out[B][S]
// (unfused gates, downstream[B][O]).  Returns 0.
int cpu_tfq_run(int mode, int64_t B, const int32_t* circ, const int32_t* gates_i,
                const uint64_t* gates_m, const float* mats, const int32_t* terms_i,
                const float* terms_c, const int32_t* paulis, int64_t O, int64_t S,
                const float* downstream, int threads, float* exp_out, float* grad_out) {
  const c64* M = reinterpret_cast<const c64*>(mats);
  std::atomic<int64_t> next(0);
  auto gate_at = [&](int64_t k) {
    Gate g;
    const int32_t* a = gates_i + 9 * k;
    g.nt = a[0]; g.b0 = a[1]; g.b1 = a[2]; g.mat = a[3]; g.nderiv = a[4];
    g.dmat0 = a[5]; g.dmat1 = a[6]; g.col0 = a[7]; g.col1 = a[8];
    g.cmask = gates_m[2 * k]; g.cval = gates_m[2 * k + 1];
    return g;
  };
  auto work = [&]() {
    std::vector<c64> psi, lam, scratch;
    while (true) {
      const int64_t i = next.fetch_add(1);
      if (i >= B) break;
      const int32_t* c = circ + 8 * i;
      const int n = c[0];
      const int64_t row = c[7];
      const uint64_t N = uint64_t(1) << n;
      psi.assign(N, c64(0, 0));
      psi[0] = c64(1, 0);
      const int gb = mode == 0 ? c[3] : c[1], ge = mode == 0 ? c[4] : c[2];
      c64 U[16];
      for (int k = gb; k < ge; ++k) {
        const Gate g = gate_at(k);
        load_matrix(M + g.mat, 1 << g.nt, false, U);
        apply(psi.data(), n, g, U);
      }
      if (mode == 0) {
        for (int64_t j = 0; j < O; ++j) exp_out[row * O + j] = 0.f;
        std::vector<double> acc(size_t(O), 0.0);
        for (int t = c[5]; t < c[6]; ++t) {
          const int32_t* ti = terms_i + 3 * t;
          const int np = ti[2] - ti[1];
          if (np == 0) {
            acc[size_t(ti[0])] += terms_c[t];
            continue;
          }
          scratch = psi;
          apply_paulis(scratch.data(), n, paulis + 2 * ti[1], np);
          acc[size_t(ti[0])] += double(terms_c[t]) * real_inner(psi.data(), scratch.data(), N);
        }
        for (int64_t j = 0; j < O; ++j) exp_out[row * O + j] = float(acc[size_t(j)]);
        continue;
      }
      // Adjoint.
      lam.assign(N, c64(0, 0));
      for (int t = c[5]; t < c[6]; ++t) {
        const int32_t* ti = terms_i + 3 * t;
        const float w = downstream[row * O + ti[0]] * terms_c[t];
        if (std::fabs(w) < 1e-5f) continue;
        scratch = psi;
        apply_paulis(scratch.data(), n, paulis + 2 * ti[1], ti[2] - ti[1]);
        for (uint64_t x = 0; x < N; ++x) lam[x] += w * scratch[x];
      }
      std::vector<double> grad(size_t(S), 0.0);
      for (int k = ge - 1; k >= gb; --k) {
        const Gate g = gate_at(k);
        load_matrix(M + g.mat, 1 << g.nt, true, U);
        apply(psi.data(), n, g, U);
        for (int d = 0; d < g.nderiv; ++d) {
          scratch = psi;
          c64 dU[16];
          load_matrix(M + (d == 0 ? g.dmat0 : g.dmat1), 1 << g.nt, false, dU);
          apply(scratch.data(), n, g, dU, /*zero_others=*/true);
          grad[size_t(d == 0 ? g.col0 : g.col1)] += 2.0 * real_inner(scratch.data(), lam.data(), N);
        }
        apply(lam.data(), n, g, U);
      }
      for (int64_t s = 0; s < S; ++s) grad_out[row * S + s] = float(grad[size_t(s)]);
    }
  };
  if (threads <= 1) {
    work();
  } else {
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t) pool.emplace_back(work);
    for (auto& th : pool) th.join();
  }
  return 0;
}

// The ComputeLarge schedule: circuits one after another, every loop over the
// state parallel over `threads` (arrays as in cpu_tfq_run).  mode 0:
// expectation (fused gates); mode 1: adjoint (unfused gates); mode 2: simulate
// the fused gates, then draw `nsamples` samples per circuit from Philox
// stream (row, 0xFFFFFFFF) into samples_out[B][nsamples] (qsim's ascending
// order).  gate_limit >= 0 bounds the work for timing a sample of a large
// circuit: modes 0 and 2 apply only the first gate_limit fused gates; mode 1
// runs only the last gate_limit gates of the reverse sweep, and fwd_limit >= 0
// applies only the first fwd_limit gates of its forward pass (lambda is
// always built in full).  Returns 0.
int cpu_tfq_run_large(int mode, int64_t B, const int32_t* circ, const int32_t* gates_i,
                      const uint64_t* gates_m, const float* mats, const int32_t* terms_i,
                      const float* terms_c, const int32_t* paulis, int64_t O, int64_t S,
                      const float* downstream, int threads, int64_t gate_limit, int64_t fwd_limit,
                      int64_t nsamples, uint64_t seed, float* exp_out, float* grad_out,
                      uint64_t* samples_out) {
  const c64* M = reinterpret_cast<const c64*>(mats);
  const int nt = std::max(1, threads);
  auto gate_at = [&](int64_t k) {
    Gate g;
    const int32_t* a = gates_i + 9 * k;
    g.nt = a[0]; g.b0 = a[1]; g.b1 = a[2]; g.mat = a[3]; g.nderiv = a[4];
    g.dmat0 = a[5]; g.dmat1 = a[6]; g.col0 = a[7]; g.col1 = a[8];
    g.cmask = gates_m[2 * k]; g.cval = gates_m[2 * k + 1];
    return g;
  };
  std::vector<c64> psi, lam, scratch;
  for (int64_t i = 0; i < B; ++i) {
    const int32_t* c = circ + 8 * i;
    const int n = c[0];
    const int64_t row = c[7];
    const int64_t N = int64_t(1) << n;
    psi.resize(size_t(N));
    if (mode != 2) scratch.resize(size_t(N));
    zero_state_par(psi.data(), N, nt);
    int gb = mode == 1 ? c[1] : c[3], ge = mode == 1 ? c[2] : c[4];
    const int64_t flim = mode == 1 ? fwd_limit : gate_limit;
    const int fe = flim >= 0 ? int(std::min<int64_t>(ge, gb + flim)) : ge;
    c64 U[16];
    for (int k = gb; k < fe; ++k) {
      const Gate g = gate_at(k);
      load_matrix(M + g.mat, 1 << g.nt, false, U);
      apply_par(psi.data(), n, g, U, false, nt);
    }
    if (mode == 2) {
      if (nsamples > 0)
        sample_par(psi.data(), N, nsamples, seed, uint32_t(row), 0xFFFFFFFFu, nt,
                   samples_out + row * nsamples);
      continue;
    }
    if (mode == 0) {
      std::vector<double> acc(size_t(O), 0.0);
      for (int t = c[5]; t < c[6]; ++t) {
        const int32_t* ti = terms_i + 3 * t;
        const int np = ti[2] - ti[1];
        if (np == 0) {
          acc[size_t(ti[0])] += terms_c[t];
          continue;
        }
        copy_par(scratch.data(), psi.data(), N, nt);
        apply_paulis_par(scratch.data(), n, paulis + 2 * ti[1], np, nt);
        acc[size_t(ti[0])] += double(terms_c[t]) * real_inner_par(psi.data(), scratch.data(), N, nt);
      }
      for (int64_t j = 0; j < O; ++j) exp_out[row * O + j] = float(acc[size_t(j)]);
      continue;
    }
    lam.resize(size_t(N));
#pragma omp parallel for num_threads(nt) schedule(static)
    for (int64_t x = 0; x < N; ++x) lam[size_t(x)] = c64(0, 0);
    for (int t = c[5]; t < c[6]; ++t) {
      const int32_t* ti = terms_i + 3 * t;
      const float w = downstream[row * O + ti[0]] * terms_c[t];
      if (std::fabs(w) < 1e-5f) continue;
      copy_par(scratch.data(), psi.data(), N, nt);
      apply_paulis_par(scratch.data(), n, paulis + 2 * ti[1], ti[2] - ti[1], nt);
      axpy_par(lam.data(), w, scratch.data(), N, nt);
    }
    std::vector<double> grad(size_t(S), 0.0);
    const int rb = gate_limit >= 0 ? std::max<int64_t>(gb, ge - gate_limit) : gb;
    for (int k = ge - 1; k >= rb; --k) {
      const Gate g = gate_at(k);
      load_matrix(M + g.mat, 1 << g.nt, true, U);
      apply_par(psi.data(), n, g, U, false, nt);
      for (int d = 0; d < g.nderiv; ++d) {
        copy_par(scratch.data(), psi.data(), N, nt);
        c64 dU[16];
        load_matrix(M + (d == 0 ? g.dmat0 : g.dmat1), 1 << g.nt, false, dU);
        apply_par(scratch.data(), n, g, dU, true, nt);
        grad[size_t(d == 0 ? g.col0 : g.col1)] +=
            2.0 * real_inner_par(scratch.data(), lam.data(), N, nt);
      }
      apply_par(lam.data(), n, g, U, false, nt);
    }
    for (int64_t s2 = 0; s2 < S; ++s2) grad_out[row * S + s2] = float(grad[size_t(s2)]);
  }
  return 0;
}

// Noisy trajectories (TFQ's ComputeSmall schedule of
// core/ops/noise/tfq_noisy_expectation.cc:284-398: trajectories spread over
// the worker threads, one float32 state each).  A gate record with nt == -1 is
// a channel: b0 = bit, nderiv = Kraus count, mat = Kraus matrices, dmat0 =
// cumulative probabilities (col1 == 1, unitary mixture; entries (cum, 0)) or
// K^dag K matrices, col0 = channel ordinal.  Channel c of trajectory t of row
// i draws Philox uniform (seed, t, (i, c)); operator k is the first with
// u < p_0 + ... + p_k.  ns[B][O]: trajectories per (row, slot).
// exp_out[B][O] = mean over trajectories + identity terms.
int cpu_tfq_noisy(int64_t B, const int32_t* circ, const int32_t* gates_i, const uint64_t* gates_m,
                  const float* mats, const int32_t* terms_i, const float* terms_c,
                  const int32_t* paulis, int64_t O, const int32_t* ns, uint64_t seed, int threads,
                  float* exp_out) {
  const c64* M = reinterpret_cast<const c64*>(mats);
  // Work items: (circuit, trajectory).
  std::vector<int64_t> first(size_t(B) + 1, 0);
  for (int64_t i = 0; i < B; ++i) {
    int T = 0;
    for (int64_t j = 0; j < O; ++j) T = std::max(T, ns[i * O + j]);
    first[size_t(i) + 1] = first[size_t(i)] + T;
  }
  const int64_t total = first[size_t(B)];
  const int nt = std::max(1, threads);
  std::vector<std::vector<double>> acc(size_t(nt), std::vector<double>(size_t(B * O), 0.0));
  std::atomic<int64_t> next(0);
  auto work = [&](int tid) {
    std::vector<c64> psi, scratch;
    std::vector<double>& my = acc[size_t(tid)];
    while (true) {
      const int64_t w = next.fetch_add(1);
      if (w >= total) break;
      int64_t i = 0;
      while (first[size_t(i) + 1] <= w) ++i;
      const int64_t t = w - first[size_t(i)];
      const int32_t* c = circ + 8 * i;
      const int n = c[0];
      const int64_t row = c[7];
      const uint64_t N = uint64_t(1) << n;
      psi.assign(N, c64(0, 0));
      psi[0] = c64(1, 0);
      c64 U[16];
      for (int k = c[1]; k < c[2]; ++k) {
        const int32_t* a = gates_i + 9 * k;
        if (a[0] >= 0) {
          Gate g;
          g.nt = a[0]; g.b0 = a[1]; g.b1 = a[2]; g.mat = a[3];
          g.cmask = gates_m[2 * k]; g.cval = gates_m[2 * k + 1];
          load_matrix(M + g.mat, 1 << g.nt, false, U);
          apply(psi.data(), n, g, U);
          continue;
        }
        const int b = a[1], nk = a[4];
        const double u = philox_u(seed, uint64_t(t), uint32_t(row), uint32_t(a[7]));
        int pick = nk - 1;
        double pp = 1.0;
        if (a[8] == 1) {
          for (int q = nk - 2; q >= 0; --q)
            if (u < double(M[a[5] + q].real())) pick = q;
        } else {
          double s00 = 0, s11 = 0, sr = 0, si = 0;
          const uint64_t bit = uint64_t(1) << b;
          for (uint64_t h = 0; h < N / 2; ++h) {
            const uint64_t x0 = ins0(h, b);
            const c64 a0 = psi[x0], a1 = psi[x0 | bit];
            s00 += std::norm(a0);
            s11 += std::norm(a1);
            sr += double(a0.real()) * a1.real() + double(a0.imag()) * a1.imag();
            si += double(a0.real()) * a1.imag() - double(a0.imag()) * a1.real();
          }
          double cum = 0;
          int last = 0;
          pick = -1;
          for (int q = 0; q < nk; ++q) {
            const c64* E = M + a[5] + 4 * q;
            const double pk = E[0].real() * s00 + E[3].real() * s11 +
                              2.0 * (double(E[1].real()) * sr - double(E[1].imag()) * si);
            if (pk > 0) last = q;
            cum += pk;
            if (pick < 0 && u < cum) {
              pick = q;
              pp = pk;
            }
          }
          if (pick < 0) {
            pick = last;
            const c64* E = M + a[5] + 4 * last;
            pp = E[0].real() * s00 + E[3].real() * s11 + 2.0 * (double(E[1].real()) * sr - double(E[1].imag()) * si);
          }
        }
        Gate g{};
        g.nt = 1;
        g.b0 = b;
        const float sc = a[8] == 1 ? 1.0f : float(1.0 / std::sqrt(pp));
        for (int e = 0; e < 4; ++e) U[e] = M[a[3] + 4 * pick + e] * sc;
        apply(psi.data(), n, g, U);
      }
      for (int tt = c[5]; tt < c[6]; ++tt) {
        const int32_t* ti = terms_i + 3 * tt;
        const int np = ti[2] - ti[1];
        if (np == 0 || t >= ns[row * O + ti[0]]) continue;
        scratch = psi;
        apply_paulis(scratch.data(), n, paulis + 2 * ti[1], np);
        my[size_t(row * O + ti[0])] += double(terms_c[tt]) * real_inner(psi.data(), scratch.data(), N);
      }
    }
  };
  if (nt <= 1) {
    work(0);
  } else {
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; ++t) pool.emplace_back(work, t);
    for (auto& th : pool) th.join();
  }
  for (int64_t i = 0; i < B; ++i) {
    const int32_t* c = circ + 8 * i;
    const int64_t row = c[7];
    for (int64_t j = 0; j < O; ++j) {
      double s = 0;
      for (int t = 0; t < nt; ++t) s += acc[size_t(t)][size_t(row * O + j)];
      double ident = 0;
      for (int tt = c[5]; tt < c[6]; ++tt)
        if (terms_i[3 * tt] == j && terms_i[3 * tt + 2] == terms_i[3 * tt + 1]) ident += terms_c[tt];
      exp_out[row * O + j] = float(s / double(ns[row * O + j]) + ident);
    }
  }
  return 0;
}

}  // extern "C"
