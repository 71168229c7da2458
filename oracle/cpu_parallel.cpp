// cpu_parallel.cpp -- C++ restatement of the reference's parallel CPU path,
// the CPU baseline timed beside the GPU.  TEST/BENCH INFRASTRUCTURE ONLY.
//
// Restates src/QuantumMatter.Infrastructure/ParallelCpuStateVectorBackend.cs
// (contiguous pair ranges per worker, one join per gate :183-237, dispatch
// :246-334, optional <=32-gate same-target blocking on 512-pair tiles with a
// 1024-amplitude scratch for strides above 512 :55-90,113-181),
// Avx512StateVectorKernel.cs (4 complex per vector for uncontrolled targets >= q2,
// complex product as v*re + permute(v)*(-+im) :109-128, diagonal skips unit
// halves :83-107), AvxStateVectorKernel.cs (2 complex per vector, the q0
// lane-permute path and the diagonal's unit-half skip :14-160),
// SimdHadamardKernel.cs (:12-63) and the scalar fallback of
// CpuStateVectorBackend.cs, chosen by the reference's dispatch (:246-276).
// Coefficients come from qmo_prepare (PreparedStateVectorGate.cs:13-70).
// Built with -ffp-contract=off, like the oracle: no fused multiply-add.
#include <immintrin.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

#include "qm_oracle.h"

namespace {

struct cx {
  double re, im;
};

inline cx cmul(cx a, cx b) { return {a.re * b.re - a.im * b.im, a.im * b.re + a.re * b.im}; }
inline cx cadd(cx a, cx b) { return {a.re + b.re, a.im + b.im}; }

// A fixed pool of workers running one range function per "Parallel.For", joined per call.
class Pool {
 public:
  explicit Pool(int workers) : n_(workers) {
    for (int i = 1; i < n_; i++) threads_.emplace_back([this, i] { loop(i); });
  }
  ~Pool() {
    {
      std::lock_guard<std::mutex> g(m_);
      stop_ = true;
      gen_++;
    }
    cv_.notify_all();
    for (auto &t : threads_) t.join();
  }
  void run(const std::function<void(int)> &fn) {
    if (n_ == 1) {
      fn(0);
      return;
    }
    {
      std::lock_guard<std::mutex> g(m_);
      fn_ = &fn;
      pending_ = n_ - 1;
      gen_++;
    }
    cv_.notify_all();
    fn(0);
    std::unique_lock<std::mutex> g(m_);
    done_.wait(g, [this] { return pending_ == 0; });
  }
  int size() const { return n_; }

 private:
  void loop(int id) {
    uint64_t seen = 0;
    for (;;) {
      const std::function<void(int)> *fn;
      {
        std::unique_lock<std::mutex> g(m_);
        cv_.wait(g, [&] { return gen_ != seen; });
        seen = gen_;
        if (stop_) return;
        fn = fn_;
      }
      (*fn)(id);
      {
        std::lock_guard<std::mutex> g(m_);
        if (--pending_ == 0) done_.notify_one();
      }
    }
  }
  int n_;
  std::vector<std::thread> threads_;
  std::mutex m_;
  std::condition_variable cv_, done_;
  const std::function<void(int)> *fn_ = nullptr;
  int pending_ = 0;
  uint64_t gen_ = 0;
  bool stop_ = false;
};

struct Gate {
  int kernel, target;
  int64_t mask;
  int swap_control;
  cx a, b, c, d;
};

Gate from_prepared(const qmo_prepared &p) {
  Gate g;
  g.kernel = p.kernel;
  g.target = p.target;
  g.mask = p.control_mask;
  g.swap_control = p.swap_control;
  g.a = {p.coeff[0], p.coeff[1]};
  g.b = {p.coeff[2], p.coeff[3]};
  g.c = {p.coeff[4], p.coeff[5]};
  g.d = {p.coeff[6], p.coeff[7]};
  return g;
}

// Avx512StateVectorKernel.Coefficient: v*re + permute(v, 0x55)*(-im, +im, ...).
#define QM_AVX512 __attribute__((target("avx512f")))
struct Coef {
  __m512d re, im;
  bool real;
  QM_AVX512 explicit Coef(cx v) {
    re = _mm512_set1_pd(v.re);
    im = _mm512_set_pd(v.im, -v.im, v.im, -v.im, v.im, -v.im, v.im, -v.im);
    real = v.im == 0;
  }
  QM_AVX512 __m512d mul(__m512d v) const {
    if (real) return _mm512_mul_pd(v, re);
    return _mm512_add_pd(_mm512_mul_pd(v, re), _mm512_mul_pd(_mm512_permute_pd(v, 0x55), im));
  }
};

QM_AVX512 void avx512_apply(cx *s, const Gate &g, int64_t start, int64_t end) {
  const int64_t bit = (int64_t)1 << g.target;
  if (g.kernel == QMO_K_DIAGONAL) {
    for (int64_t i = start; i < end;) {
      int64_t run = std::min(std::min(bit - (i & (bit - 1)), end - i), (int64_t)4096);
      cx ph = (i & bit) == 0 ? g.a : g.d;
      if (!(ph.re == 1 && ph.im == 0)) {
        Coef c(ph);
        int64_t k = 0;
        double *p = (double *)(s + i);
        for (; k + 4 <= run; k += 4) _mm512_storeu_pd(p + 2 * k, c.mul(_mm512_loadu_pd(p + 2 * k)));
        for (; k < run; k++) s[i + k] = cmul(s[i + k], ph);
      }
      i += run;
    }
    return;
  }
  Coef a(g.a), b(g.b), c(g.c), d(g.d);
  const __m512d h = _mm512_set1_pd(g.a.re);
  for (int64_t pair = start; pair < end;) {
    int64_t low = pair & (bit - 1);
    int64_t first = ((pair - low) << 1) | low, second = first | bit;
    int64_t run = std::min(std::min(bit - low, end - pair), (int64_t)4096);
    double *z = (double *)(s + first), *o = (double *)(s + second);
    int64_t k = 0;
    for (; k + 4 <= run; k += 4) {
      __m512d x = _mm512_loadu_pd(z + 2 * k), y = _mm512_loadu_pd(o + 2 * k);
      if (g.kernel == QMO_K_EXCHANGE) {
        _mm512_storeu_pd(z + 2 * k, y);
        _mm512_storeu_pd(o + 2 * k, x);
      } else if (g.kernel == QMO_K_HADAMARD) {
        __m512d sx = _mm512_mul_pd(x, h), sy = _mm512_mul_pd(y, h);
        _mm512_storeu_pd(z + 2 * k, _mm512_add_pd(sx, sy));
        _mm512_storeu_pd(o + 2 * k, _mm512_sub_pd(sx, sy));
      } else {
        _mm512_storeu_pd(z + 2 * k, _mm512_add_pd(a.mul(x), b.mul(y)));
        _mm512_storeu_pd(o + 2 * k, _mm512_add_pd(c.mul(x), d.mul(y)));
      }
    }
    for (; k < run; k++) {
      cx x = s[first + k], y = s[second + k];
      if (g.kernel == QMO_K_EXCHANGE) {
        s[first + k] = y;
        s[second + k] = x;
      } else {
        s[first + k] = cadd(cmul(g.a, x), cmul(g.b, y));
        s[second + k] = cadd(cmul(g.c, x), cmul(g.d, y));
      }
    }
    pair += run;
  }
}

// AvxStateVectorKernel.Coefficient (AvxStateVectorKernel.cs:144-160): two
// complex values per 256-bit vector; v*re + permute(v, 0x05)*(-im1, im1, -im2, im2),
// or v*re alone when both coefficients are real.
#define QM_AVX __attribute__((target("avx")))
struct Coef2 {
  __m256d re, im;
  bool real;
  QM_AVX Coef2(cx first, cx second) {
    re = _mm256_set_pd(second.re, second.re, first.re, first.re);
    im = _mm256_set_pd(second.im, -second.im, first.im, -first.im);
    real = first.im == 0 && second.im == 0;
  }
  QM_AVX explicit Coef2(cx v) : Coef2(v, v) {}
  QM_AVX __m256d mul(__m256d v) const {
    if (real) return _mm256_mul_pd(v, re);
    return _mm256_add_pd(_mm256_mul_pd(v, re), _mm256_mul_pd(_mm256_permute_pd(v, 0x05), im));
  }
};

// AvxStateVectorKernel.ApplyDiagonal (:106-141): odd range boundaries one by
// one, runs of up to 4096 (even) amplitudes, identity halves skipped above q0.
QM_AVX void avx_diagonal(cx *s, const Gate &g, int64_t start, int64_t end) {
  const int64_t bit = (int64_t)1 << g.target;
  for (int64_t i = start; i < end;) {
    if ((i & 1) != 0 || i + 1 == end) {
      s[i] = cmul(s[i], (i & bit) == 0 ? g.a : g.d);
      i++;
      continue;
    }
    const cx ph = (i & bit) == 0 ? g.a : g.d;
    int64_t run = std::min(std::min(bit == 1 ? end - i : bit - (i & (bit - 1)), end - i), (int64_t)4096);
    run &= ~(int64_t)1;
    if (bit != 1 && ph.re == 1 && ph.im == 0) {
      i += run;
      continue;
    }
    const Coef2 c = bit == 1 ? Coef2(g.a, g.d) : Coef2(ph);
    double *p = (double *)(s + i);
    for (int64_t lane = 0; lane < 2 * run; lane += 4) _mm256_storeu_pd(p + lane, c.mul(_mm256_loadu_pd(p + lane)));
    i += run;
  }
}

// AvxStateVectorKernel.Apply (:14-103): for q0 one vector holds both amplitudes
// of a pair (half swap for an exchange, (A x + B y, C x + D y) from broadcast
// halves otherwise); above q0 two pairs per vector, a scalar tail for odd runs.
QM_AVX void avx_apply(cx *s, const Gate &g, int64_t start, int64_t end) {
  if (g.kernel == QMO_K_DIAGONAL) {
    avx_diagonal(s, g, start, end);
    return;
  }
  const int64_t bit = (int64_t)1 << g.target;
  const Coef2 a(g.a), b(g.b), c(g.c), d(g.d), ac(g.a, g.c), bd(g.b, g.d);
  const bool exchange = g.kernel == QMO_K_EXCHANGE;
  for (int64_t pair = start; pair < end;) {
    const int64_t low = pair & (bit - 1);
    const int64_t first = ((pair - low) << 1) | low, second = first | bit;
    if (bit == 1) {
      const int64_t count = std::min((int64_t)4096, end - pair);
      double *p = (double *)(s + first);
      for (int64_t off = 0; off < 4 * count; off += 4) {
        const __m256d in = _mm256_loadu_pd(p + off);
        __m256d out;
        if (exchange) {
          out = _mm256_permute2f128_pd(in, in, 0x01);
        } else {
          const __m256d x = _mm256_permute2f128_pd(in, in, 0x00), y = _mm256_permute2f128_pd(in, in, 0x11);
          out = _mm256_add_pd(ac.mul(x), bd.mul(y));
        }
        _mm256_storeu_pd(p + off, out);
      }
      pair += count;
      continue;
    }
    const int64_t run = std::min(std::min(bit - low, end - pair), (int64_t)4096);
    double *z = (double *)(s + first), *o = (double *)(s + second);
    int64_t lane = 0;
    for (; lane + 4 <= 2 * run; lane += 4) {
      const __m256d x = _mm256_loadu_pd(z + lane), y = _mm256_loadu_pd(o + lane);
      if (exchange) {
        _mm256_storeu_pd(z + lane, y);
        _mm256_storeu_pd(o + lane, x);
      } else {
        _mm256_storeu_pd(z + lane, _mm256_add_pd(a.mul(x), b.mul(y)));
        _mm256_storeu_pd(o + lane, _mm256_add_pd(c.mul(x), d.mul(y)));
      }
    }
    if (lane < 2 * run) {
      const int64_t k = lane / 2;
      const cx x = s[first + k], y = s[second + k];
      if (exchange) {
        s[first + k] = y;
        s[second + k] = x;
      } else {
        s[first + k] = cadd(cmul(g.a, x), cmul(g.b, y));
        s[second + k] = cadd(cmul(g.c, x), cmul(g.d, y));
      }
    }
    pair += run;
  }
}

// SimdHadamardKernel.Apply (SimdHadamardKernel.cs:12-63) with .NET's 256-bit
// Vector<double> (2 complex per vector): (h x + h y, h x - h y), scalar tail.
QM_AVX void simd_hadamard(cx *s, int target, int64_t start, int64_t end) {
  const int64_t bit = (int64_t)1 << target;
  const double h = std::sqrt(0.5);
  const __m256d scale = _mm256_set1_pd(h);
  for (int64_t pair = start; pair < end;) {
    const int64_t low = pair & (bit - 1);
    const int64_t first = ((pair - low) << 1) | low, second = first | bit;
    const int64_t run = std::min(bit - low, end - pair);
    double *z = (double *)(s + first), *o = (double *)(s + second);
    int64_t off = 0;
    while (off < run) {
      if (off + 2 <= run) {
        const __m256d sx = _mm256_mul_pd(_mm256_loadu_pd(z + 2 * off), scale);
        const __m256d sy = _mm256_mul_pd(_mm256_loadu_pd(o + 2 * off), scale);
        _mm256_storeu_pd(z + 2 * off, _mm256_add_pd(sx, sy));
        _mm256_storeu_pd(o + 2 * off, _mm256_sub_pd(sx, sy));
        off += 2;
      } else {
        const cx x = s[first + off], y = s[second + off];
        s[first + off] = {h * x.re + h * y.re, h * x.im + h * y.im};
        s[second + off] = {h * x.re - h * y.re, h * x.im - h * y.im};
        off++;
      }
    }
    pair += run;
  }
}

// ParallelCpuStateVectorBackend.ApplyRange scalar fallback (:278-333).
void scalar_apply(cx *s, const Gate &g, int64_t start, int64_t end) {
  const int64_t bit = (int64_t)1 << g.target;
  if (g.kernel == QMO_K_DIAGONAL) {
    for (int64_t i = start; i < end; i++)
      if ((i & g.mask) == g.mask) s[i] = cmul(s[i], (i & bit) == 0 ? g.a : g.d);
    return;
  }
  const int64_t sc = g.kernel == QMO_K_SWAP ? (int64_t)1 << g.swap_control : 0;
  for (int64_t pair = start; pair < end; pair++) {
    int64_t low = pair & (bit - 1);
    int64_t first = ((pair - low) << 1) | low;
    if ((first & g.mask) != g.mask) continue;
    int64_t second = first | bit;
    if (g.kernel == QMO_K_SWAP) {
      if ((first & sc) == 0) continue;
      second ^= sc;
      std::swap(s[first], s[second]);
    } else if (g.kernel == QMO_K_EXCHANGE) {
      std::swap(s[first], s[second]);
    } else {
      cx x = s[first], y = s[second];
      s[first] = cadd(cmul(g.a, x), cmul(g.b, y));
      s[second] = cadd(cmul(g.c, x), cmul(g.d, y));
    }
  }
}

bool has_avx512() { return __builtin_cpu_supports("avx512f"); }
bool has_avx() { return __builtin_cpu_supports("avx"); }

// ParallelCpuStateVectorBackend.ApplyRange dispatch (:246-276): AVX-512 for
// uncontrolled targets >= q2; otherwise AVX for uncontrolled matrix, exchange
// and diagonal gates (and H on q0); the SIMD Hadamard for higher H targets;
// scalar loops for controlled gates and swaps.
void apply_range(cx *s, const Gate &g, int64_t start, int64_t end, bool simd, bool avx512) {
  const int64_t bit = (int64_t)1 << g.target;
  const bool kinds = g.kernel == QMO_K_MATRIX || g.kernel == QMO_K_HADAMARD || g.kernel == QMO_K_EXCHANGE ||
                     g.kernel == QMO_K_DIAGONAL;
  if (simd && avx512 && g.mask == 0 && bit >= 4 && has_avx512() && kinds) {
    avx512_apply(s, g, start, end);
    return;
  }
  if (simd && g.mask == 0 && has_avx() && kinds && (g.kernel != QMO_K_HADAMARD || bit == 1)) {
    avx_apply(s, g, start, end);
    return;
  }
  if (g.kernel == QMO_K_HADAMARD && g.mask == 0 && simd && has_avx() && bit >= 2) {
    simd_hadamard(s, g.target, start, end);
    return;
  }
  scalar_apply(s, g, start, end);
}

bool can_block(const Gate &g) {
  return g.mask == 0 && (g.kernel == QMO_K_MATRIX || g.kernel == QMO_K_HADAMARD ||
                         g.kernel == QMO_K_EXCHANGE || g.kernel == QMO_K_DIAGONAL);
}

// ApplyBlockedRange (:113-181).
void blocked_range(cx *s, const Gate *gates, int count, int64_t start, int64_t end, bool simd,
                   bool avx512) {
  const int64_t bit = (int64_t)1 << gates[0].target;
  cx scratch[1024];
  for (int64_t pair = start; pair < end;) {
    int64_t low = pair & (bit - 1);
    int64_t first = ((pair - low) << 1) | low;
    int64_t run = std::min((int64_t)512, end - pair);
    bool contiguous = low == 0 && bit <= run;
    if (contiguous)
      run -= run % bit;
    else
      run = std::min(run, bit - low);
    if (bit > 512) {
      int local = 63 - __builtin_clzll((uint64_t)run);
      run = (int64_t)1 << local;
      int64_t second = first | bit;
      std::memcpy(scratch, s + first, sizeof(cx) * run);
      std::memcpy(scratch + run, s + second, sizeof(cx) * run);
      for (int k = 0; k < count; k++) {
        Gate lg = gates[k];
        lg.target = local;
        int64_t items = lg.kernel == QMO_K_DIAGONAL ? 2 * run : run;
        apply_range(scratch, lg, 0, items, simd, avx512);
      }
      std::memcpy(s + first, scratch, sizeof(cx) * run);
      std::memcpy(s + second, scratch + run, sizeof(cx) * run);
      pair += run;
      continue;
    }
    for (int k = 0; k < count; k++) {
      const Gate &g = gates[k];
      if (g.kernel != QMO_K_DIAGONAL) {
        apply_range(s, g, pair, pair + run, simd, avx512);
      } else if (contiguous) {
        apply_range(s, g, first, first + 2 * run, simd, avx512);
      } else {
        apply_range(s, g, first, first + run, simd, avx512);
        int64_t second = first | bit;
        apply_range(s, g, second, second + run, simd, avx512);
      }
    }
    pair += run;
  }
}

}  // namespace

// ParallelCpuOptions.MinimumParallelQubits (ParallelCpuOptions.cs): widths below
// it run on one worker; the reference's default is 16, its accelerator
// benchmark forces 1 (AcceleratorBenchmarks.cs:35).
extern "C" int qmo_parallel_execute_ex(int qubits, const qmo_op *ops, int64_t n_ops, double *amps,
                                       int workers, int simd, int avx512, int blocking,
                                       int min_parallel_qubits) {
  if (qubits < 1 || qubits > 34 || !amps || workers < 1 || workers > 256) return QMO_ERR_ARG;
  cx *s = (cx *)amps;
  const int64_t len = (int64_t)1 << qubits;
  const bool parallel = qubits >= min_parallel_qubits && workers > 1;
  Pool pool(parallel ? workers : 1);
  std::vector<Gate> gates;
  for (int64_t i = 0; i < n_ops; i++) {
    qmo_prepared p;
    int rc = qmo_prepare(&ops[i], &p);
    if (rc) return rc;
    gates.push_back(from_prepared(p));
  }
  for (size_t k = 0; k < gates.size(); k++) {
    const Gate &g = gates[k];
    if (g.kernel == QMO_K_IDENTITY) continue;
    if (blocking && can_block(g)) {
      int count = 1;
      while (count < 32 && k + count < gates.size() && gates[k + count].target == g.target &&
             can_block(gates[k + count]))
        count++;
      if (count > 1) {
        const int64_t pairs = len / 2;
        const int w = parallel ? (int)std::min<int64_t>(pairs, pool.size()) : 1;
        const Gate *gp = &gates[k];
        pool.run([&](int id) {
          if (id >= w) return;
          int64_t a = pairs * id / w, b = pairs * (id + 1) / w;
          blocked_range(s, gp, count, a, b, simd, avx512);
        });
        k += count - 1;
        continue;
      }
    }
    const int64_t items = g.kernel == QMO_K_DIAGONAL ? len : len / 2;
    const int w = parallel ? (int)std::min<int64_t>(items, pool.size()) : 1;
    pool.run([&](int id) {
      if (id >= w) return;
      int64_t a = items * id / w, b = items * (id + 1) / w;
      apply_range(s, g, a, b, simd, avx512);
    });
  }
  return QMO_OK;
}

extern "C" int qmo_parallel_execute(int qubits, const qmo_op *ops, int64_t n_ops, double *amps,
                                    int workers, int simd, int avx512, int blocking) {
  return qmo_parallel_execute_ex(qubits, ops, n_ops, amps, workers, simd, avx512, blocking, 16);
}
