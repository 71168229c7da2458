// Compiled fused passes: each planned tile pass becomes a straight-line HIP
// kernel (hipRTC, gfx950), with every register index, control mask, LDS
// address and load/store offset a compile-time constant and only the gate
// coefficients read from memory.  This removes the per-op dispatch of the
// interpreting kernel (fused.hip): no runtime switch over register bits, so
// no register shuffling between ops, and register controls select at compile
// time.  Kernels are cached by their source text: the pass structure and, for
// each uncontrolled rotation taken in factored form, which of its two forms
// its angle selects (the larger of |cos| and |sin| of half the angle is taken
// out), not the coefficients; so re-running a circuit structure with new
// angles reuses the kernels of every pass whose rotations keep their forms,
// and compiles the others.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>
#include <dlfcn.h>
#include <spawn.h>
#include <sys/wait.h>
#include <unistd.h>

#include <cerrno>
#include <cmath>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "jit.hpp"
#include "rtc.hpp"
#include "qm_internal.hpp"

namespace qm {
namespace {

// Device helpers of the generated kernels.  Every update writes each output
// over the input it replaces after that input's last use, so the compiler
// needs no copies; register controls are template masks.
const char *kJitPrelude = R"QMJIT(
struct __attribute__((aligned(16))) qj_c128 { double re, im; };
struct __attribute__((aligned(8))) qj_c64 { float re, im; };
__device__ __forceinline__ double qj_fma(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float qj_fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

template <typename F, int N, int B, int M>
__device__ __forceinline__ void qj_rx(F (&vr)[N], F (&vi)[N], F c, F s) {
#pragma unroll
  for (int r = 0; r < N; r++) {
    if ((r & (1 << B)) || (r & M) != M) continue;
    const int q = r | (1 << B);
    const F a = c * vr[r], b = s * vr[r], e = c * vi[r], f = s * vi[r];
    vr[r] = qj_fma(s, vi[q], a);
    vi[q] = qj_fma(c, vi[q], -b);
    vi[r] = qj_fma(-s, vr[q], e);
    vr[q] = qj_fma(c, vr[q], f);
  }
}
template <typename F, int N, int B, int M>
__device__ __forceinline__ void qj_ry(F (&vr)[N], F (&vi)[N], F c, F s) {
#pragma unroll
  for (int r = 0; r < N; r++) {
    if ((r & (1 << B)) || (r & M) != M) continue;
    const int q = r | (1 << B);
    const F a = c * vr[r], b = s * vr[r], e = c * vi[r], f = s * vi[r];
    vr[r] = qj_fma(-s, vr[q], a);
    vr[q] = qj_fma(c, vr[q], b);
    vi[r] = qj_fma(-s, vi[q], e);
    vi[q] = qj_fma(c, vi[q], f);
  }
}
template <typename F, int N, int B, int M>
__device__ __forceinline__ void qj_h(F (&vr)[N], F (&vi)[N], F h) {
#pragma unroll
  for (int r = 0; r < N; r++) {
    if ((r & (1 << B)) || (r & M) != M) continue;
    const int q = r | (1 << B);
    const F a = h * vr[r], e = h * vi[r];
    vr[r] = qj_fma(h, vr[q], a);
    vr[q] = qj_fma(-h, vr[q], a);
    vi[r] = qj_fma(h, vi[q], e);
    vi[q] = qj_fma(-h, vi[q], e);
  }
}
template <typename F, int N, int B, int M>
__device__ __forceinline__ void qj_y(F (&vr)[N], F (&vi)[N]) {
#pragma unroll
  for (int r = 0; r < N; r++) {
    if ((r & (1 << B)) || (r & M) != M) continue;
    const int q = r | (1 << B);
    const F x0 = vr[r], x1 = vi[r];
    vr[r] = vi[q];
    vi[r] = -vr[q];
    vr[q] = -x1;
    vi[q] = x0;
  }
}
template <typename F, int N, int B, int M>
__device__ __forceinline__ void qj_x(F (&vr)[N], F (&vi)[N]) {
#pragma unroll
  for (int r = 0; r < N; r++) {
    if ((r & (1 << B)) || (r & M) != M) continue;
    const int q = r | (1 << B);
    const F a = vr[r], b = vi[r];
    vr[r] = vr[q];
    vi[r] = vi[q];
    vr[q] = a;
    vi[q] = b;
  }
}
// Address space of the coefficient table when it is read through scalar loads.
#ifndef QJ_CONST
#define QJ_CONST __attribute__((address_space(4)))
#endif
template <typename F, int N, int B, int M, typename P>
__device__ __forceinline__ void qj_mat(F (&vr)[N], F (&vi)[N], P c) {
  const F c0 = (F)c[0], c1 = (F)c[1], c2 = (F)c[2], c3 = (F)c[3];
  const F c4 = (F)c[4], c5 = (F)c[5], c6 = (F)c[6], c7 = (F)c[7];
#pragma unroll
  for (int r = 0; r < N; r++) {
    if ((r & (1 << B)) || (r & M) != M) continue;
    const int q = r | (1 << B);
    const F x0 = vr[r], x1 = vi[r], y0 = vr[q], y1 = vi[q];
    vr[r] = qj_fma(-c3, y1, qj_fma(c2, y0, qj_fma(-c1, x1, c0 * x0)));
    vi[r] = qj_fma(c2, y1, qj_fma(c3, y0, qj_fma(c0, x1, c1 * x0)));
    vr[q] = qj_fma(-c7, y1, qj_fma(c6, y0, qj_fma(-c5, x1, c4 * x0)));
    vi[q] = qj_fma(c6, y1, qj_fma(c7, y0, qj_fma(c4, x1, c5 * x0)));
  }
}
// Uncontrolled rotations with a common factor taken out (applied once per
// pass): Rx = c (I - i t X) with t = s/c (form a, |c| >= |s|) or s (u I - i X)
// with u = c/s (form b); Ry likewise with -i X replaced by the real -iY.
// One fused multiply-add per output.
template <typename F, int N, int B>
__device__ __forceinline__ void qj_rxa(F (&vr)[N], F (&vi)[N], F t) {
#pragma unroll
  for (int r = 0; r < N; r++) {
    if (r & (1 << B)) continue;
    const int q = r | (1 << B);
    const F xr = vr[r], xi = vi[r], yr = vr[q], yi = vi[q];
    vr[r] = qj_fma(t, yi, xr);
    vi[r] = qj_fma(-t, yr, xi);
    vr[q] = qj_fma(t, xi, yr);
    vi[q] = qj_fma(-t, xr, yi);
  }
}
template <typename F, int N, int B>
__device__ __forceinline__ void qj_rxb(F (&vr)[N], F (&vi)[N], F u) {
#pragma unroll
  for (int r = 0; r < N; r++) {
    if (r & (1 << B)) continue;
    const int q = r | (1 << B);
    const F xr = vr[r], xi = vi[r], yr = vr[q], yi = vi[q];
    vr[r] = qj_fma(u, xr, yi);
    vi[r] = qj_fma(u, xi, -yr);
    vr[q] = qj_fma(u, yr, xi);
    vi[q] = qj_fma(u, yi, -xr);
  }
}
template <typename F, int N, int B>
__device__ __forceinline__ void qj_rya(F (&vr)[N], F (&vi)[N], F t) {
#pragma unroll
  for (int r = 0; r < N; r++) {
    if (r & (1 << B)) continue;
    const int q = r | (1 << B);
    const F xr = vr[r], xi = vi[r], yr = vr[q], yi = vi[q];
    vr[r] = qj_fma(-t, yr, xr);
    vi[r] = qj_fma(-t, yi, xi);
    vr[q] = qj_fma(t, xr, yr);
    vi[q] = qj_fma(t, xi, yi);
  }
}
template <typename F, int N, int B>
__device__ __forceinline__ void qj_ryb(F (&vr)[N], F (&vi)[N], F u) {
#pragma unroll
  for (int r = 0; r < N; r++) {
    if (r & (1 << B)) continue;
    const int q = r | (1 << B);
    const F xr = vr[r], xi = vi[r], yr = vr[q], yi = vi[q];
    vr[r] = qj_fma(u, xr, -yr);
    vi[r] = qj_fma(u, xi, -yi);
    vr[q] = qj_fma(u, yr, xr);
    vi[q] = qj_fma(u, yi, xi);
  }
}
template <typename F>
__device__ __forceinline__ void qj_mul(F &xr, F &xi, F pr, F pi) {
  const F a = xi * pi, b = xr * pi;
  xr = qj_fma(xr, pr, -a);
  xi = qj_fma(xi, pr, b);
}
// diag(A, D) on register bit B (A skipped unless FULL), registers under mask M.
template <typename F, int N, int B, int M, bool FULL>
__device__ __forceinline__ void qj_diag(F (&vr)[N], F (&vi)[N], F ar, F ai, F dr, F di) {
#pragma unroll
  for (int r = 0; r < N; r++) {
    if ((r & M) != M) continue;
    if (r & (1 << B))
      qj_mul(vr[r], vi[r], dr, di);
    else if (FULL)
      qj_mul(vr[r], vi[r], ar, ai);
  }
}
template <typename F, int N, int M>
__device__ __forceinline__ void qj_phase(F (&vr)[N], F (&vi)[N], F pr, F pi) {
#pragma unroll
  for (int r = 0; r < N; r++) {
    if ((r & M) != M) continue;
    qj_mul(vr[r], vi[r], pr, pi);
  }
}
)QMJIT";

std::string hex(uint64_t v) {
  char b[32];
  std::snprintf(b, sizeof b, "0x%llxull", (unsigned long long)v);
  return b;
}

// Source of one pass kernel, and the coefficient layout it reads.
// Code-generation variant (tuning knob QM_JIT_VARIANT; default 11665 = bits 0,
// 4, 7, 8, 10, 11 and 13 in complex128, 12177 = those and bit 9 in complex64):
// bit 0 exchanges through an LDS buffer of reals (half the LDS,
// real then imaginary parts, so more workgroups fit a CU), bit 2 forms the
// lane part of each LDS address at its transpose instead of once per kernel,
// bit 3 loads and stores the state with non-temporal hints (a pass never
// re-reads what it wrote), bit 4 takes the common factor out of uncontrolled
// rotations and register diagonals (ratio forms, one multiply-add per output)
// and applies the product of the factors once per pass (only with bit 7:
// without it, builds of some of these passes computed wrong states on the
// device, 2 to 9 of 60 layer circuits of 21..30 qubits depending on
// optimisation level and occupancy, while the same sources are exact in the
// host emulation; those builds hold ~200 coefficients in scalar registers
// spilled to vector lanes, and with bit 7, no scalar spills, 140 such
// circuits are exact), bit 5 keeps exchanges conditioned on thread or outside
// bits as per-tile flags (X frames, see pass_source), bit 6 issues the next
// tile's loads before the current tile's work (compiled for one workgroup per
// CU, so the registers of both tiles fit), bit 7 loads the coefficients inside
// the tile loop instead of once per kernel, bit 8 forms the swizzled LDS
// addresses of each exchange there instead of holding those of every exchange
// across the tile loop (which spilled up to 700 B per thread), bit 9 splits
// each swizzled LDS address into the XOR of the few distinct parts that
// overlap the lane part and an immediate offset (fewer address instructions
// and paired LDS accesses; 6 % faster in complex64, 3 % slower in complex128
// at 16384-amplitude tiles, where the extra live addresses spill), bit 10
// (with bit 0) lays out each exchange in LDS so that the lanes one LDS
// instruction serves hit distinct banks on both sides (exchange_layout; the
// XOR-fold layout of the tables left 76 % extra bank-conflict cycles; 4 %
// faster in complex128; with bit 9, 4 % faster than bit 9 alone in complex64),
// bit 11 (with bit 7) reads the coefficients through the constant address
// space, so they are scalar loads rather than vector loads whose waits also
// wait for the state stores in flight (2 % faster in complex128, 5 % in
// complex64), bit 12 (with bits 5 and 10) carries the flips of bit 5 through
// X gates controlled by one flipped register bit (the register permutation
// plus the control's flag) and folds the flips left at an exchange or at the
// store into the LDS write or store addresses instead of swapping registers
// under a branch, with each exchange's layout chosen for the folded writer
// addresses and a per-thread factor for flipped diagonals (the moves of the
// conditional swaps go, 7 % fewer VALU instructions, but 2 % slower in
// complex128 and 5 % in complex64: more spills and waits), bit 13 (with bit 4,
// without bit 5) holds factored diagonals on register bits back while the ops
// that follow do not target their bits and applies them together, one product
// per register (27 % fewer of their complex products on the 30-qubit layer
// circuit; 1-3 % faster in complex128, 4 % in complex64).
int jit_variant(int precision) {
  static const int v = [] {
    const char *e = std::getenv("QM_JIT_VARIANT");
    return e ? std::atoi(e) : -1;
  }();
  return v >= 0 ? v : precision == QM_C64 ? 12177 : 11665;
}

int variant_of(const FusedShape &sh, int precision) {
  return sh.variant >= 0 ? sh.variant : jit_variant(precision);
}

// Factored forms (bit 4) without per-tile coefficient loads (bit 7) have
// computed wrong device states for some passes of 21..30-qubit circuits while
// their sources run exactly in the host emulation; the cause is not known, so
// such a variant is never loaded onto a device.  Generating and compiling it
// without loading (jit_compile_check, the host emulation) stays possible.
void check_device_variant(int variant) {
  if ((variant & 16) && !(variant & 128))
    fail(QM_ERR_UNSUPPORTED, "code variant " + std::to_string(variant) +
                                 " (factored forms without per-tile coefficient loads) is not run on a device");
}

size_t jit_lds_bytes(int precision, const FusedShape &sh) {
  const bool split = (variant_of(sh, precision) & 1) != 0;
  return (split ? (size_t)amp_bytes(precision) / 2 : (size_t)amp_bytes(precision)) << sh.K;
}

// Rank of vectors over GF(2) (bit masks).
static int gf2_rank(std::vector<uint32_t> v) {
  int rank = 0;
  for (int bit = 31; bit >= 0; bit--) {
    size_t i = rank;
    while (i < v.size() && !((v[i] >> bit) & 1)) i++;
    if (i == v.size()) continue;
    std::swap(v[rank], v[i]);
    for (size_t k = 0; k < v.size(); k++)
      if (k != (size_t)rank && ((v[k] >> bit) & 1)) v[k] ^= v[rank];
    rank++;
  }
  return rank;
}

// LDS layout of one exchange: an invertible GF(2)-linear map S on the K tile
// bits, given by the images cols[p] of the unit vectors, under which the
// amplitude at tile position x sits at element S(x).  Writers (mapping
// `from`) and readers (mapping `to`) both address through S, so any
// invertible S is correct; it is chosen so that the lanes an LDS instruction
// serves together hit distinct banks on both sides: for the first b thread
// bits (b lanes per 128 bytes, and b + 1 per 256 bytes), the low b (b + 1)
// bits of their images are independent.  S keeps bits >= B (B = b + 1) in
// place and gives every position a B-bit low part: positions below B a basis
// of the low bits, the others any value (block triangular, so invertible).
// Values are found by a seeded randomised search; identity if none is found.
// Folded flips (position, thread-bit mask): a writer whose thread bits in the
// mask have odd parity writes the amplitude at that position flipped, so the
// image of writer thread bit k also carries the positions of the folds whose
// mask holds k.
static std::vector<uint32_t> exchange_layout(const DevPass &d, int from, int to, int K, int L,
                                             int elem_bytes,
                                             const std::vector<std::pair<int, uint32_t>> &folds = {}) {
  int b = 0;
  while ((elem_bytes << (b + 1)) <= 128) b++;
  b = std::min({b, L - 1, K - 1});
  const int B = b + 1;
  std::vector<uint32_t> cols(K);
  for (int p = 0; p < K; p++) cols[p] = 1u << p;
  if (b < 1) return cols;
  const uint32_t low = (1u << B) - 1;
  // Groups: (images as sets of positions XORed, bits that must be independent).
  std::vector<std::pair<std::vector<std::vector<int>>, int>> groups;
  for (int slot : {from, to})
    for (int n : {b, B}) {
      std::vector<std::vector<int>> g;
      for (int k = 0; k < n; k++) {
        std::vector<int> e{d.lanepos[slot][k]};
        if (slot == from)
          for (const auto &f : folds)
            if ((f.second >> k) & 1) e.push_back(f.first);
        g.push_back(e);
      }
      groups.push_back({g, n});
    }
  std::vector<int> order;
  for (auto &g : groups)
    for (const auto &e : g.first)
      for (int p : e)
        if (std::find(order.begin(), order.end(), p) == order.end()) order.push_back(p);
  uint64_t rng = 0x9E3779B97F4A7C15ull ^ ((uint64_t)from << 32) ^ (uint64_t)to;
  auto next = [&] {
    rng ^= rng << 13;
    rng ^= rng >> 7;
    rng ^= rng << 17;
    return rng;
  };
  std::vector<uint32_t> val(K, 0);
  std::vector<char> set(K, 0);
  auto ok = [&] {
    for (auto &g : groups) {
      std::vector<uint32_t> v;
      for (const auto &e : g.first) {
        bool all = true;
        uint32_t x = 0;
        for (int p : e) {
          all = all && set[p];
          x ^= val[p];
        }
        if (all) v.push_back(x & ((1u << g.second) - 1));
      }
      if (gf2_rank(v) != (int)v.size()) return false;
    }
    std::vector<uint32_t> basis;  // low parts of positions below B stay independent
    for (int p = 0; p < B && p < K; p++)
      if (set[p]) basis.push_back(val[p]);
    return gf2_rank(basis) == (int)basis.size();
  };
  for (int attempt = 0; attempt < 4000; attempt++) {
    std::fill(set.begin(), set.end(), 0);
    bool dead = false;
    for (int p : order) {
      set[p] = 1;
      bool found = false;
      const uint32_t start = (uint32_t)next();
      for (uint32_t c = 0; c < low && !found; c++) {
        val[p] = ((start + c) % low) + 1;
        found = ok();
      }
      if (!found) {
        dead = true;
        break;
      }
    }
    if (dead) continue;
    // Complete the low basis with the unused positions below B.
    for (int p = 0; p < B && p < K; p++) {
      if (set[p]) continue;
      set[p] = 1;
      bool found = false;
      for (uint32_t c = 1; c <= low && !found; c++) {
        val[p] = c;
        found = ok();
      }
      if (!found) {
        dead = true;
        break;
      }
    }
    if (dead) continue;
    for (int p = 0; p < K; p++) cols[p] = (p < B ? 0u : (1u << p)) | (set[p] ? val[p] : 0u);
    return cols;
  }
  for (int p = 0; p < K; p++) cols[p] = 1u << p;
  return cols;
}

std::string pass_source(const FusedPass &p, const FusedShape &sh, int precision,
                        std::vector<double> &coef, std::vector<int32_t> *prov = nullptr) {
  const int variant = variant_of(sh, precision);
  const bool split = (variant & 1) != 0, occ3 = (variant & 2) != 0, inline_j = (variant & 4) != 0;
  const bool nontemporal = (variant & 8) != 0, factored = (variant & 16) != 0, frames = (variant & 32) != 0;
  const bool prefetch = (variant & 64) != 0, tile_coef = (variant & 128) != 0;
  const bool launder = (variant & 256) != 0, fold = (variant & 512) != 0;
  const bool banked = (variant & 1024) != 0 && split;
  const bool scalar_coef = (variant & 2048) != 0 && tile_coef;
  const bool fold_frames = (variant & 4096) != 0 && frames && banked;
  const bool merge_diag = (variant & 8192) != 0 && !frames && factored;
  const DevPass &d = p.dev;
  const int L = sh.L, R = sh.R, N = 1 << R;
  const bool batched = sh.state_bits > 0;
  const int tile_shift = sh.state_bits - sh.K;
  if (batched && (!tile_coef || !scalar_coef || prefetch || factored || merge_diag || d.out_place || d.store_nruns))
    fail(QM_ERR_UNSUPPORTED, "batched compiled passes need per-tile scalar coefficients, no factored forms, "
                             "no prefetch and in-place passes");
  const bool c128 = precision == QM_C128;
  const char *F = c128 ? "double" : "float";
  const char *T = c128 ? "qj_c128" : "qj_c64";
  std::ostringstream o;
  // QM_OCC (workgroups per CU the register allocation must allow) is defined
  // when the kernel is compiled; the launch bound's second argument counts
  // waves per SIMD, and a workgroup of 2^L threads puts 2^L / 256 waves on
  // each of the CU's four SIMDs.
  (void)occ3;
  o << "extern \"C\" __global__ __launch_bounds__(" << (1 << L) << ", (QM_OCC * " << (1 << L)
    << " + 255) / 256) void qm_pass(";
  // A pass of a lane-layout plan reads one buffer (s) and writes the other
  // (so): its tiles store to other index bits than they load from.
  if (d.out_place)
    o << "const " << T << " *__restrict__ s, " << T << " *__restrict__ so";
  else
    o << T << " *__restrict__ s";
  o << ", const double *__restrict__ " << (tile_coef ? "cf_arg" : "cf") << ", unsigned long long tiles) {\n";
  o << "  typedef " << F << " F;\n";
  if (split)
    o << "  extern __shared__ F lds[];\n";
  else
    o << "  extern __shared__ " << T << " lds[];\n";
  o << "  const unsigned t = threadIdx.x;\n";
  auto lane_off = [&](const char *name, const uint64_t *offs) {
    o << "  const unsigned long long " << name << " = 0ull";
    for (int k = 0; k < L; k++) o << " | ((t & " << (1u << k) << "u) ? " << hex(offs[k]) << " : 0ull)";
    o << ";\n";
  };
  lane_off("ll", d.load_lane_off);
  lane_off("sl", d.store_lane_off);
  // Lane part of the swizzled LDS address of every mapping a transpose touches.
  std::vector<int> used;
  auto use = [&](int slot) {
    if (std::find(used.begin(), used.end(), slot) == used.end()) used.push_back(slot);
  };
  int cur = d.load_direct ? 0 : kIo;
  for (int ph = 0; ph < d.nphases; ph++) {
    if (cur != ph) use(cur), use(ph);
    cur = ph;
  }
  if (!d.store_direct) use(cur), use(kIoStore);
  auto lane_addr = [&](int slot) {
    std::string e = "(0u";
    for (int k = 0; k < L; k++)
      e += " ^ ((t & " + std::to_string(1u << k) + "u) ? " + std::to_string(d.lane_swz[slot][k]) + "u : 0u)";
    return e + ")";
  };
  if (!inline_j)
    for (int slot : used) o << "  const unsigned j" << slot << " = " << lane_addr(slot) << ";\n";
  auto base_expr = [&](const std::string &tv) {
    std::string e = "0ull";
    for (int k = 0; k < d.nruns; k++)
      e += " | (((" + tv + " >> " + std::to_string((int)d.run_src[k]) + ") & " + hex((1ull << d.run_len[k]) - 1) +
           ") << " + std::to_string((int)d.run_dst[k]) + ")";
    return e;
  };
  auto loads = [&](const char *re, const char *im, const std::string &b) {
    o << "    { const " << T << " *src = " << (batched ? "sp" : "s") << " + ((" << b << ") | ll);\n";
    for (int r = 0; r < N; r++) {
      if (nontemporal)
        o << "      { const " << T << " *a = src + " << hex(d.load_reg_off[r]) << "; " << re << "[" << r
          << "] = __builtin_nontemporal_load(&a->re); " << im << "[" << r << "] = __builtin_nontemporal_load(&a->im); }\n";
      else
        o << "      { const " << T << " a = src[" << hex(d.load_reg_off[r]) << "]; " << re << "[" << r << "] = a.re; "
          << im << "[" << r << "] = a.im; }\n";
    }
    o << "    }\n";
  };
  if (prefetch) {
    // Software pipelining: the next tile's loads are issued before this tile's
    // work, so they are in flight during its exchanges and arithmetic.
    o << "  F nr[" << N << "], ni[" << N << "];\n";
    o << "  if (blockIdx.x < tiles) {\n";
    loads("nr", "ni", base_expr("(unsigned long long)blockIdx.x"));
    o << "  }\n";
  }
  o << "  for (unsigned long long tau = blockIdx.x; tau < tiles; tau += gridDim.x) {\n";
  if (tile_coef) {
    // The coefficient pointer is re-derived every tile (an empty asm that
    // "changes" it), so coefficient loads stay next to their uses instead of
    // being hoisted out of the loop and held in (spilled) scalar registers.
    if (scalar_coef) {
      // The laundered address is cast to the constant address space, so each
      // coefficient is a scalar load (waited on with lgkmcnt) instead of a
      // vector load, whose wait would also wait for the previous tile's
      // state stores still in flight.
      if (batched)
        o << "    unsigned long long cfa = (unsigned long long)(cf_arg + (tau >> " << tile_shift
          << ") * @NCOEF@ull);\n";
      else
        o << "    unsigned long long cfa = (unsigned long long)cf_arg;\n";
      o << "    asm volatile(\"\" : \"+s\"(cfa));\n";
      o << "    const QJ_CONST double *cf = (const QJ_CONST double *)cfa;\n";
    } else {
      o << "    const double *cf = cf_arg;\n";
      o << "    asm volatile(\"\" : \"+s\"(cf));\n";
    }
  }
  if (batched) {
    // Batched passes: the tile index picks the state, then the tile in it.
    o << "    " << T << " *const sp = s + ((tau >> " << tile_shift << ") << " << sh.state_bits << ");\n";
    o << "    const unsigned long long base = " << base_expr("(tau & " + hex((1ull << tile_shift) - 1) + ")")
      << ";\n";
  } else {
    o << "    const unsigned long long base = " << base_expr("tau") << ";\n";
  }
  if (d.store_nruns) {
    std::string e = "0ull";
    for (int k = 0; k < d.store_nruns; k++)
      e += " | (((tau >> " + std::to_string((int)d.store_run_src[k]) + ") & " +
           hex((1ull << d.store_run_len[k]) - 1) + ") << " + std::to_string((int)d.store_run_dst[k]) + ")";
    o << "    const unsigned long long sbase = " << e << ";\n";
  }
  o << "    F vr[" << N << "], vi[" << N << "];\n";
  if (prefetch) {
    for (int r = 0; r < N; r++) o << "    vr[" << r << "] = nr[" << r << "]; vi[" << r << "] = ni[" << r << "];\n";
    o << "    if (tau + gridDim.x < tiles) {\n";
    loads("nr", "ni", base_expr("(tau + gridDim.x)"));
    o << "    }\n";
  } else {
    loads("vr", "vi", "base");
  }
  // A flip folded into an exchange: its position and, when its condition is
  // linear, the thread bits whose parity it is.
  struct Fold {
    int lp;
    uint32_t tmask;
    bool linear;
  };
  auto transpose = [&](int from, int to, const std::vector<Fold> &folds) {
    o << "    {\n";
    if (inline_j) {
      o << "    const unsigned j" << from << " = " << lane_addr(from) << ";\n";
      o << "    const unsigned j" << to << " = " << lane_addr(to) << ";\n";
    }
    // With launder, the lane addresses are re-derived (an empty asm that
    // "changes" them) before each exchange, so the 2 x N swizzled LDS
    // addresses are computed there instead of being hoisted out of the tile
    // loop and held, for every exchange of the pass at once, in (spilled)
    // registers.
    if (banked) {
      // Per-exchange layout (exchange_layout): lane parts formed here from a
      // laundered thread index, register parts as constants.
      std::vector<std::pair<int, uint32_t>> fl;
      for (const Fold &f : folds)
        if (f.linear) fl.push_back({f.lp, f.tmask});
      const std::vector<uint32_t> cols = exchange_layout(d, from, to, sh.K, L, c128 ? 8 : 4, fl);
      auto S = [&](uint32_t x) {
        uint32_t y = 0;
        for (int q = 0; q < sh.K; q++)
          if ((x >> q) & 1) y ^= cols[q];
        return y;
      };
      auto lane_expr = [&](int slot) {
        std::string e = "(0u";
        for (int k = 0; k < L; k++)
          e += " ^ ((tt & " + std::to_string(1u << k) + "u) ? " + std::to_string(S(1u << d.lanepos[slot][k])) +
               "u : 0u)";
        return e + ")";
      };
      // With fold, the register constant's bits outside every lane image are
      // added (immediate offset) instead of XORed, as in `at` below.
      // Folded frames: the writer XORs its lane part with the image of each
      // flipped position, so those bits are XOR parts of the write addresses.
      uint32_t fmask = 0;
      std::string fterm;
      for (const Fold &f : folds) {
        fmask |= S(1u << f.lp);
        fterm += " ^ (fq" + std::to_string(f.lp) + " ? " + std::to_string(S(1u << f.lp)) + "u : 0u)";
      }
      auto addr = [&](const char *j, int slot, uint32_t c) {
        if (!fold) return std::string("lds[") + j + " ^ " + std::to_string(c) + "u]";
        uint32_t m = slot == from ? fmask : 0u;
        for (int k = 0; k < L; k++) m |= S(1u << d.lanepos[slot][k]);
        return std::string("lds[(") + j + " ^ " + std::to_string(c & m) + "u) + " + std::to_string(c & ~m) + "u]";
      };
      o << "    unsigned tt = t;\n    asm volatile(\"\" : \"+v\"(tt));\n";
      o << "    const unsigned jw = " << lane_expr(from) << fterm << ", jr = " << lane_expr(to) << ";\n";
      for (const char *part : {"vr", "vi"}) {
        o << "    __syncthreads();\n";
        for (int r = 0; r < N; r++)
          o << "    " << addr("jw", from, S(d.regloc[from][r])) << " = " << part << "[" << r << "];\n";
        o << "    __syncthreads();\n";
        for (int r = 0; r < N; r++)
          o << "    " << part << "[" << r << "] = " << addr("jr", to, S(d.regloc[to][r])) << ";\n";
      }
      o << "    }\n";
      return;
    }
    const std::string jw = launder ? "jw" : "j" + std::to_string(from);
    const std::string jr = launder ? "jr" : "j" + std::to_string(to);
    // Address of register r under a mapping: the lane part j XOR the
    // register's constant c.  With fold, the bits of c outside every bit j can
    // hold are added instead (same value, as j has none of them), so they
    // become the LDS instruction's immediate offset, and only the few
    // distinct XOR parts are computed.
    auto at = [&](const std::string &j, int slot, unsigned c) {
      if (!fold) return "lds[" + j + " ^ " + std::to_string(c) + "u]";
      unsigned m = 0;
      for (int k = 0; k < L; k++) m |= d.lane_swz[slot][k];
      return "lds[(" + j + " ^ " + std::to_string(c & m) + "u) + " + std::to_string(c & ~m) + "u]";
    };
    if (split) {
      for (const char *part : {"vr", "vi"}) {
        o << "    __syncthreads();\n";
        if (launder) {
          o << "    {\n    unsigned jw = j" << from << ", jr = j" << to << ";\n";
          o << "    asm volatile(\"\" : \"+v\"(jw), \"+v\"(jr));\n";
        }
        for (int r = 0; r < N; r++)
          o << "    " << at(jw, from, d.regswz[from][r]) << " = " << part << "[" << r << "];\n";
        o << "    __syncthreads();\n";
        for (int r = 0; r < N; r++)
          o << "    " << part << "[" << r << "] = " << at(jr, to, d.regswz[to][r]) << ";\n";
        if (launder) o << "    }\n";
      }
    } else {
      o << "    __syncthreads();\n";
      for (int r = 0; r < N; r++)
        o << "    { " << T << " a; a.re = vr[" << r << "]; a.im = vi[" << r << "]; "
          << at("j" + std::to_string(from), from, d.regswz[from][r]) << " = a; }\n";
      o << "    __syncthreads();\n";
      for (int r = 0; r < N; r++)
        o << "    { const " << T << " a = " << at("j" + std::to_string(to), to, d.regswz[to][r]) << "; vr[" << r
          << "] = a.re; vi[" << r << "] = a.im; }\n";
    }
    o << "    }\n";
  };
  auto cfix = [&](const double *v, int n) {
    const size_t at = coef.size();
    coef.insert(coef.end(), v, v + n);
    if (prov) {
      // Batched passes: every slot is a copy of an op coefficient.
      const char *b0 = (const char *)p.ops.data(), *x = (const char *)v;
      if (x < b0 || x >= b0 + p.ops.size() * sizeof(TileOp))
        fail(QM_ERR_UNSUPPORTED, "batched compiled passes need coefficients copied from the ops");
      const size_t byte = (size_t)(x - b0), op = byte / sizeof(TileOp);
      const size_t off = (byte % sizeof(TileOp) - offsetof(TileOp, c)) / sizeof(double);
      for (int i = 0; i < n; i++) {
        prov->push_back((int32_t)op);
        prov->push_back((int32_t)(off + (size_t)i));
      }
    }
    return at;
  };
  const std::string tmpl = std::string("<F, ") + std::to_string(N) + ", ";
  cur = d.load_direct ? 0 : kIo;
  const TileOp *ops = p.ops.data();
  const int K = sh.K;
  auto lp_reg = [&](int slot, int bit) { return __builtin_ctz((unsigned)d.regloc[slot][1u << bit]); };
  auto fname = [](int lp) { return "fq" + std::to_string(lp); };
  // X frames: an exchange on register bit b conditioned on lane or outside bits
  // only is not performed on the registers; it sets a per-tile flag fq<lp> for
  // its tile position.  While the flag is set, the data of that position is
  // held with the bit flipped: conditions on it are inverted, Ry and diagonal
  // coefficients are conjugated by X, and the store address flips the bit.  A
  // register control, or an H, Y or general matrix on it, performs the pending
  // exchange first.  A flag set under thread-bit conditions describes the data
  // each thread holds, so it stays valid across an LDS transpose only while its
  // condition positions keep their thread bits and its position stays a
  // register bit; otherwise the exchange is performed before the transpose.
  struct Frame {
    bool lane = false;                     // depends on thread bits
    std::vector<std::pair<int, int>> deps;  // (tile position, thread bit)
    uint32_t tmask = 0;                    // the condition's thread bits (XOR), when linear
    bool nonlinear = false;                // a product of thread bits
  };
  std::vector<char> framed((size_t)K, 0);
  std::vector<Frame> fdep((size_t)K);
  auto is_source = [&](const TileOp &op) {
    return frames && op.kind == G_X && op.rbit >= 0 && op.reg_cmask == 0 && (op.lane_cmask || op.out_cmask);
  };
  std::vector<char> declared((size_t)K, 0);
  for (int ph = 0; ph < d.nphases; ph++)
    for (int i = 0; i < d.phase_op_count[ph]; i++) {
      const TileOp &op = ops[d.phase_op_begin[ph] + i];
      if (is_source(op)) declared[(size_t)lp_reg(ph, op.rbit)] = 1;
    }
  if (fold_frames)
    for (int ph = 0; ph < d.nphases; ph++)
      for (int i = 0; i < d.phase_op_count[ph]; i++) {
        const TileOp &op = ops[d.phase_op_begin[ph] + i];
        if (op.kind == G_X && op.rbit >= 0) declared[(size_t)lp_reg(ph, op.rbit)] = 1;
      }
  for (int lp = 0; lp < K; lp++)
    if (declared[(size_t)lp]) o << "    bool " << fname(lp) << " = false;\n";
  auto cond = [&](const TileOp &op, int slot) {
    std::vector<std::string> parts;
    uint32_t plain = 0;
    for (int k = 0; k < L; k++)
      if (op.lane_cmask & (1u << k)) {
        const int lp = d.lanepos[slot][k];
        if (framed[(size_t)lp])
          parts.push_back("((((t >> " + std::to_string(k) + ") & 1u) != 0u) != " + fname(lp) + ")");
        else
          plain |= 1u << k;
      }
    if (plain) parts.insert(parts.begin(), "(t & " + std::to_string(plain) + "u) == " + std::to_string(plain) + "u");
    if (op.out_cmask) parts.push_back("(base & " + hex(op.out_cmask) + ") == " + hex(op.out_cmask));
    std::string c;
    for (const std::string &x : parts) c += (c.empty() ? "" : " && ") + x;
    return c;
  };
  auto hi_expr = [&](const TileOp &op, int slot) {
    std::string hi;
    if (op.dlane) {
      const int lp = d.lanepos[slot][__builtin_ctz(op.dlane)];
      hi = "((t & " + std::to_string(op.dlane) + "u) != 0u)";
      if (framed[(size_t)lp]) hi = "(" + hi + " != " + fname(lp) + ")";
    }
    if (op.dout) hi += std::string(hi.empty() ? "" : " || ") + "(base & " + hex(op.dout) + ") != 0ull";
    return hi.empty() ? std::string("false") : hi;
  };
  // Performs the pending exchange of register bit `bit` of mapping `slot`.
  auto settle = [&](int slot, int bit) {
    const int lp = lp_reg(slot, bit);
    if (!framed[(size_t)lp]) return;
    o << "    if (" << fname(lp) << ") { qj_x" << tmpl << bit << ", 0>(vr, vi); } " << fname(lp) << " = false;\n";
    framed[(size_t)lp] = 0;
    fdep[(size_t)lp] = Frame{};
  };
  auto reg_bit = [&](int slot, int lp) {
    for (int b = 0; b < R; b++)
      if (lp_reg(slot, b) == lp) return b;
    return -1;
  };
  // LDS transpose, first performing the exchanges whose flags would not
  // survive it.
  // With fold_frames, such an exchange is instead folded into the LDS write
  // addresses: the amplitude a thread holds at a flipped position is written
  // where it belongs, so no register moves.
  auto move_to = [&](int from, int to) {
    std::vector<Fold> folds;
    for (int lp = 0; lp < K; lp++) {
      if (!framed[(size_t)lp] || !fdep[(size_t)lp].lane) continue;
      bool keep = reg_bit(to, lp) >= 0;
      for (const auto &dk : fdep[(size_t)lp].deps) keep = keep && d.lanepos[to][dk.second] == dk.first;
      if (keep) continue;
      if (fold_frames)
        folds.push_back({lp, fdep[(size_t)lp].tmask, !fdep[(size_t)lp].nonlinear});
      else
        settle(from, reg_bit(from, lp));
    }
    transpose(from, to, folds);
    for (const Fold &f : folds) {
      o << "    " << fname(f.lp) << " = false;\n";
      framed[(size_t)f.lp] = 0;
      fdep[(size_t)f.lp] = Frame{};
    }
  };
  auto sel = [&](int lp, const std::string &when_set, const std::string &otherwise) {
    return framed[(size_t)lp] ? "(" + fname(lp) + " ? " + when_set + " : " + otherwise + ")" : otherwise;
  };
  // Factored forms: uncontrolled ops only (their factor is a scalar multiple of
  // the whole state, so it commutes with everything and is applied at the end
  // of the last phase).
  auto uncontrolled = [](const TileOp &op) { return !op.reg_cmask && !op.lane_cmask && !op.out_cmask; };
  auto factor_rot = [&](const TileOp &op) {
    return factored && op.rbit >= 0 && (op.kind == G_RX || op.kind == G_RY) && uncontrolled(op);
  };
  auto factor_diag = [&](const TileOp &op) {
    return factored && op.rbit < 0 && op.kind == G_DIAG && op.dreg && uncontrolled(op);
  };
  bool any_factor = false;
  for (int ph = 0; ph < d.nphases; ph++)
    for (int i = 0; i < d.phase_op_count[ph]; i++) {
      const TileOp &op = ops[d.phase_op_begin[ph] + i];
      any_factor |= factor_rot(op) || factor_diag(op);
    }
  double gre = 1, gim = 0;  // product of the factors taken out
  size_t gslot = 0;
  const bool use_gneg = any_factor && std::find(declared.begin(), declared.end(), 1) != declared.end();
  if (any_factor) {
    const double zero[2] = {0, 0};
    gslot = cfix(zero, 2);
  }
  if (use_gneg) o << "    bool gneg = false;\n";
  auto cf_str = [](size_t k) { return "(F)cf[" + std::to_string(k) + "]"; };
  // Per-thread factor of a phase (fold_frames, phases that end with a
  // multiplication of every register anyway: merged diagonals or the pass's
  // factors): a factored diagonal on a flipped position multiplies the other
  // half by the inverse ratio and the thread's factor by the ratio, and a
  // factored Ry in form b on a flipped position negates the thread's factor,
  // instead of multiplying both halves.
  // Merged register diagonals (merge_diag): a factored diagonal on register
  // bit b (multiply the registers with bit b set by q) commutes with every op
  // that does not target b as a pair, so it is held back and combined with
  // the other held diagonals; when an op targets a held bit, or the phase
  // ends, each register is multiplied once by the product of the held ratios
  // of its set bits (2^k - 1 distinct products for k held bits, formed on the
  // host): (2^k - 1) N / 2^k complex products instead of k N / 2.
  std::vector<std::pair<int, std::pair<double, double>>> held;  // (register bit, ratio)
  auto flush_held = [&]() {
    if (held.empty()) return;
    const int k = (int)held.size();
    std::vector<size_t> slot((size_t)1 << k, 0);
    for (int m = 1; m < (1 << k); m++) {
      double pr = 1, pi = 0;
      for (int i = 0; i < k; i++)
        if ((m >> i) & 1) {
          const double ar = held[(size_t)i].second.first, ai = held[(size_t)i].second.second;
          const double nr = pr * ar - pi * ai, ni = pr * ai + pi * ar;
          pr = nr;
          pi = ni;
        }
      const double v[2] = {pr, pi};
      slot[(size_t)m] = cfix(v, 2);
    }
    o << "    {";
    for (int r = 0; r < N; r++) {
      int m = 0;
      for (int i = 0; i < k; i++)
        if ((r >> held[(size_t)i].first) & 1) m |= 1 << i;
      if (!m) continue;
      o << " qj_mul(vr[" << r << "], vi[" << r << "], " << cf_str(slot[(size_t)m]) << ", "
        << cf_str(slot[(size_t)m] + 1) << ");";
    }
    o << " }\n";
    held.clear();
  };
  for (int ph = 0; ph < d.nphases; ph++) {
    if (cur != ph) move_to(cur, ph);
    cur = ph;
    const bool gl_ok = fold_frames && (d.phase_diag_count[ph] > 0 || ph == d.nphases - 1);
    const std::string glr = "glr" + std::to_string(ph), gli = "gli" + std::to_string(ph);
    bool gl_used = false;
    if (gl_ok) o << "    F " << glr << " = (F)1, " << gli << " = (F)0;\n";
    for (int i = 0; i < d.phase_op_count[ph]; i++) {
      const TileOp &op = ops[d.phase_op_begin[ph] + i];
      if (merge_diag) {
        if (factor_diag(op)) {
          // diag(A, D) = A diag(1, D/A): A joins the pass's factor, D/A is held.
          const double ar = op.c[0], ai = op.c[1], dr = op.c[6], di = op.c[7];
          const double den = ar * ar + ai * ai;
          const double qr = (dr * ar + di * ai) / den, qi = (di * ar - dr * ai) / den;
          const double nr = gre * ar - gim * ai, ni = gre * ai + gim * ar;
          gre = nr;
          gim = ni;
          const int b = __builtin_ctz(op.dreg);
          bool merged = false;
          for (auto &h : held)
            if (h.first == b) {
              const double hr = h.second.first * qr - h.second.second * qi;
              const double hi = h.second.first * qi + h.second.second * qr;
              h.second = {hr, hi};
              merged = true;
            }
          if (!merged) held.push_back({b, {qr, qi}});
          continue;
        }
        if (op.rbit >= 0)
          for (const auto &h : held)
            if (h.first == op.rbit) {
              flush_held();
              break;
            }
      }
      // An X on register bit b controlled by one register bit c whose
      // position is flipped under condition f: X_b^(r_c xor f) is the
      // register permutation X_b^(r_c) followed by the flip X_b^f, so the
      // permutation is done on the registers and f is added to b's flag.
      if (fold_frames && op.kind == G_X && op.rbit >= 0 && __builtin_popcount(op.reg_cmask) == 1 &&
          !op.lane_cmask && !op.out_cmask && framed[(size_t)lp_reg(ph, __builtin_ctz(op.reg_cmask))]) {
        const int clp = lp_reg(ph, __builtin_ctz(op.reg_cmask)), xlp = lp_reg(ph, op.rbit);
        o << "    { qj_x" << tmpl << op.rbit << ", " << op.reg_cmask << ">(vr, vi); }\n";
        o << "    " << fname(xlp) << " ^= " << fname(clp) << ";\n";
        const Frame g = fdep[(size_t)clp];
        Frame &f = fdep[(size_t)xlp];
        f.lane = f.lane || g.lane;
        f.deps.insert(f.deps.end(), g.deps.begin(), g.deps.end());
        f.tmask ^= g.tmask;
        f.nonlinear = f.nonlinear || g.nonlinear;
        framed[(size_t)xlp] = 1;
        continue;
      }
      for (int b = 0; b < R; b++)
        if (op.reg_cmask & (1u << b)) settle(ph, b);
      const int tlp = op.rbit >= 0 ? lp_reg(ph, op.rbit) : (op.dreg ? lp_reg(ph, __builtin_ctz(op.dreg)) : -1);
      if (op.rbit >= 0 && (op.kind == G_H || op.kind == G_Y || op.kind == G_MAT)) settle(ph, op.rbit);
      // A factored Ry in form b on a position flipped under thread-bit
      // conditions: its factor's sign would depend on the thread's data, so
      // the flip is performed first, or (fold_frames) the Ry is applied
      // unfactored with its sine's sign chosen per thread.
      bool unfactor = false;
      if (factor_rot(op) && op.kind == G_RY && framed[(size_t)tlp] && fdep[(size_t)tlp].lane &&
          std::fabs(op.c[0]) < std::fabs(op.c[1])) {
        if (gl_ok)
          ;
        else if (fold_frames)
          unfactor = true;
        else
          settle(ph, op.rbit);
      }
      if (is_source(op)) {
        o << "    " << fname(tlp) << " ^= (" << cond(op, ph) << ");\n";
        Frame &f = fdep[(size_t)tlp];
        if (__builtin_popcount(op.lane_cmask) == 1 && !framed[(size_t)d.lanepos[ph][__builtin_ctz(op.lane_cmask)]])
          f.tmask ^= op.lane_cmask;
        else if (op.lane_cmask)
          f.nonlinear = true;
        for (int k = 0; k < L; k++)
          if (op.lane_cmask & (1u << k)) {
            const int c = d.lanepos[ph][k];
            f.lane = true;
            f.deps.push_back({c, k});
            if (framed[(size_t)c]) {
              f.lane = f.lane || fdep[(size_t)c].lane;
              f.deps.insert(f.deps.end(), fdep[(size_t)c].deps.begin(), fdep[(size_t)c].deps.end());
            }
          }
        framed[(size_t)tlp] = 1;
        continue;
      }
      const std::string c = cond(op, ph);
      o << "    " << (c.empty() ? "{" : "if (" + c + ") {") << " ";
      const std::string m = std::to_string(op.reg_cmask);
      if (factor_rot(op) && !unfactor) {
        const double c = op.c[0], sn = op.c[1];
        const bool form_a = std::fabs(c) >= std::fabs(sn);
        const double ratio[1] = {form_a ? sn / c : c / sn};
        const double f = form_a ? c : sn;
        gre *= f;
        gim *= f;
        const size_t k = cfix(ratio, 1);
        std::string arg = cf_str(k);
        if (op.kind == G_RY) {
          // X Ry(theta) X = Ry(-theta): the ratio changes sign, and in form b
          // so does the factor taken out.
          arg = sel(tlp, "-" + cf_str(k), cf_str(k));
          if (!form_a && framed[(size_t)tlp] && fdep[(size_t)tlp].lane) {
            const std::string f = fname(tlp);
            o << glr << " = " << f << " ? -" << glr << " : " << glr << "; " << gli << " = " << f << " ? -" << gli
              << " : " << gli << "; ";
            gl_used = true;
          } else if (!form_a && framed[(size_t)tlp]) {
            o << "gneg ^= " << fname(tlp) << "; ";
          }
        }
        o << (op.kind == G_RX ? "qj_rx" : "qj_ry") << (form_a ? "a" : "b") << tmpl << op.rbit << ">(vr, vi, " << arg
          << ");";
      } else if (factor_diag(op)) {
        // diag(A, D) = A diag(1, D/A), |A| = 1 for the reference's diagonal kinds
        // but divided exactly as complex numbers in general.
        const double ar = op.c[0], ai = op.c[1], dr = op.c[6], di = op.c[7];
        const double den = ar * ar + ai * ai;
        const double q[2] = {(dr * ar + di * ai) / den, (di * ar - dr * ai) / den};
        const double nr = gre * ar - gim * ai, ni = gre * ai + gim * ar;
        gre = nr;
        gim = ni;
        const size_t k = cfix(q, 2);
        const int b = __builtin_ctz(op.dreg);
        const double dden = dr * dr + di * di;
        if (framed[(size_t)tlp] && gl_ok && dden > 0) {
          const double qi[2] = {(ar * dr + ai * di) / dden, (ai * dr - ar * di) / dden};
          const size_t k2 = cfix(qi, 2);
          const std::string f = fname(tlp);
          o << "qj_diag" << tmpl << b << ", 0, false>(vr, vi, (F)1, (F)0, (" << f << " ? " << cf_str(k2) << " : "
            << cf_str(k) << "), (" << f << " ? " << cf_str(k2 + 1) << " : " << cf_str(k + 1) << ")); ";
          o << "const F fr = " << f << " ? " << cf_str(k) << " : (F)1, fi = " << f << " ? " << cf_str(k + 1)
            << " : (F)0; const F nr = " << glr << " * fr - " << gli << " * fi, ni = " << glr << " * fi + " << gli
            << " * fr; " << glr << " = nr; " << gli << " = ni;";
          gl_used = true;
        } else if (framed[(size_t)tlp])
          o << "qj_diag" << tmpl << b << ", 0, true>(vr, vi, " << sel(tlp, cf_str(k), "(F)1") << ", "
            << sel(tlp, cf_str(k + 1), "(F)0") << ", " << sel(tlp, "(F)1", cf_str(k)) << ", "
            << sel(tlp, "(F)0", cf_str(k + 1)) << ");";
        else
          o << "qj_diag" << tmpl << b << ", 0, false>(vr, vi, (F)1, (F)0, " << cf_str(k) << ", " << cf_str(k + 1)
            << ");";
      } else if (op.rbit >= 0) {
        const std::string bm = std::to_string(op.rbit) + ", " + m + ">";
        switch (op.kind) {
          case G_RX: {
            const size_t k = cfix(op.c, 2);
            o << "qj_rx" << tmpl << bm << "(vr, vi, " << cf_str(k) << ", " << cf_str(k + 1) << ");";
            break;
          }
          case G_RY: {
            const size_t k = cfix(op.c, 2);
            o << "qj_ry" << tmpl << bm << "(vr, vi, " << cf_str(k) << ", "
              << sel(tlp, "-" + cf_str(k + 1), cf_str(k + 1)) << ");";
            break;
          }
          case G_H: {
            const size_t k = cfix(op.c, 1);
            o << "qj_h" << tmpl << bm << "(vr, vi, " << cf_str(k) << ");";
            break;
          }
          case G_Y:
            o << "qj_y" << tmpl << bm << "(vr, vi);";
            break;
          case G_X:
            o << "qj_x" << tmpl << bm << "(vr, vi);";
            break;
          default: {  // G_MAT
            const size_t k = cfix(op.c, 8);
            o << "qj_mat" << tmpl << bm << "(vr, vi, cf + " << k << ");";
            break;
          }
        }
      } else {
        const bool full = op.kind == G_DIAG;
        const size_t k = cfix(op.c, 8);
        const std::string ar = full ? cf_str(k) : "(F)1", ai = full ? cf_str(k + 1) : "(F)0";
        const std::string dr = cf_str(k + 6), di = cf_str(k + 7);
        if (op.dreg) {
          const int b = __builtin_ctz(op.dreg);
          if (framed[(size_t)tlp])
            o << "qj_diag" << tmpl << b << ", " << m << ", true>(vr, vi, " << sel(tlp, dr, ar) << ", "
              << sel(tlp, di, ai) << ", " << sel(tlp, ar, dr) << ", " << sel(tlp, ai, di) << ");";
          else
            o << "qj_diag" << tmpl << b << ", " << m << ", " << (full ? "true" : "false") << ">(vr, vi, " << ar
              << ", " << ai << ", " << dr << ", " << di << ");";
        } else {
          const std::string hi = hi_expr(op, ph);
          if (!full)
            o << "if (" << hi << ") qj_phase" << tmpl << m << ">(vr, vi, " << dr << ", " << di << ");";
          else
            o << "const bool hi = " << hi << "; qj_phase" << tmpl << m << ">(vr, vi, hi ? " << dr << " : " << ar
              << ", hi ? " << di << " : " << ai << ");";
        }
      }
      o << " }\n";
    }
    flush_held();
    const int dcnt = d.phase_diag_count[ph];
    const bool last = ph == d.nphases - 1;
    std::string g0 = "1", g1 = "0";
    if (any_factor && last) {
      g0 = cf_str(gslot), g1 = cf_str(gslot + 1);
      if (use_gneg) g0 = "(gneg ? -" + g0 + " : " + g0 + ")", g1 = "(gneg ? -" + g1 + " : " + g1 + ")";
    }
    if (!dcnt && any_factor && last && !gl_used)
      o << "    qj_phase" << tmpl << "0>(vr, vi, " << g0 << ", " << g1 << ");\n";
    if (!dcnt && gl_used)
      o << "    { const F pr = " << g0 << " * " << glr << " - " << g1 << " * " << gli << ", pi = " << g0 << " * " << gli
        << " + " << g1 << " * " << glr << "; qj_phase" << tmpl << "0>(vr, vi, pr, pi); }\n";
    if (dcnt) {
      // Merged lane/outside diagonals: one per-lane phase, one multiplication
      // (the last phase's starts from the factors taken out of the pass).
      o << "    { F pr = " << g0 << ", pi = " << g1 << ";\n";
      if (gl_used)
        o << "      { const F nr = pr * " << glr << " - pi * " << gli << ", ni = pr * " << gli << " + pi * " << glr
          << "; pr = nr; pi = ni; }\n";
      for (int i = 0; i < dcnt; i++) {
        const TileOp &q = ops[d.phase_diag_begin[ph] + i];
        const size_t k = cfix(q.c, 8);
        const bool full = q.kind == G_DIAG;
        std::string on = "true";
        if (q.lane_cmask || q.out_cmask) on = cond(q, ph);
        o << "      { const bool on = " << on << "; const bool hi = " << hi_expr(q, ph) << ";\n";
        o << "        const F fr = on ? (hi ? " << cf_str(k + 6) << " : " << (full ? cf_str(k) : "(F)1")
          << ") : (F)1;\n";
        o << "        const F fi = on ? (hi ? " << cf_str(k + 7) << " : " << (full ? cf_str(k + 1) : "(F)0")
          << ") : (F)0;\n";
        o << "        const F nr = pr * fr - pi * fi, ni = pr * fi + pi * fr; pr = nr; pi = ni; }\n";
      }
      o << "      qj_phase" << tmpl << "0>(vr, vi, pr, pi); }\n";
    }
  }
  if (any_factor) {
    coef[gslot] = gre;
    coef[gslot + 1] = gim;
  }
  if (!d.store_direct) move_to(cur, kIoStore);
  // Pending exchanges at the store: a thread bit flips the store address, a
  // register bit is performed on the registers.
  // With fold_frames a flipped register bit b moves the address instead: the
  // registers with bit b clear store that many amplitudes higher, those with
  // it set lower (one base address per combination of flipped bits).
  const int st = d.store_direct ? cur : kIoStore;
  std::string flips;
  std::vector<std::pair<int, int>> rflips;  // (register bit, position)
  for (int lp = 0; lp < K; lp++) {
    if (!framed[(size_t)lp]) continue;
    int lane = -1;
    for (int k = 0; k < L; k++)
      if (d.lanepos[st][k] == lp) lane = k;
    if (lane >= 0) {
      flips += " ^ (" + fname(lp) + " ? " + hex(1ull << d.store_lane_gbit[lane]) + " : 0ull)";
    } else {
      for (int b = 0; b < R; b++)
        if (lp_reg(st, b) == lp) {
          if (fold_frames)
            rflips.push_back({b, lp});
          else
            settle(st, b);
        }
    }
  }
  o << "    { " << T << " *dst = " << (batched ? "sp" : d.out_place ? "so" : "s") << " + ("
    << (d.store_nruns ? "sbase" : "base") << " | (sl" << flips << "));\n";
  const int nf = (int)rflips.size();
  for (int m = 0; nf && m < (1 << nf); m++) {
    o << "      " << T << " *const dst" << m << " = dst";
    for (int i = 0; i < nf; i++)
      o << " + (" << fname(rflips[(size_t)i].second) << " ? " << (((m >> i) & 1) ? "-" : "") << "(long long)"
        << hex(d.store_reg_off[1u << rflips[(size_t)i].first]) << " : 0ll)";
    o << ";\n";
  }
  auto dst_of = [&](int r) {
    if (!nf) return std::string("dst");
    int m = 0;
    for (int i = 0; i < nf; i++)
      if ((r >> rflips[(size_t)i].first) & 1) m |= 1 << i;
    return "dst" + std::to_string(m);
  };
  for (int r = 0; r < N; r++) {
    if (nontemporal)
      o << "      { " << T << " *a = " << dst_of(r) << " + " << hex(d.store_reg_off[r])
        << "; __builtin_nontemporal_store(vr[" << r << "], &a->re); __builtin_nontemporal_store(vi[" << r
        << "], &a->im); }\n";
    else
      o << "      { " << T << " a; a.re = vr[" << r << "]; a.im = vi[" << r << "]; " << dst_of(r) << "["
        << hex(d.store_reg_off[r]) << "] = a; }\n";
  }
  o << "    }\n  }\n}\n";
  if (!batched) return o.str();
  std::string out = o.str();
  const std::string mark = "@NCOEF@";
  for (size_t at = out.find(mark); at != std::string::npos; at = out.find(mark, at))
    out.replace(at, mark.size(), std::to_string(coef.size()));
  return out;
}

struct Compiled {
  hipModule_t mod = nullptr;
  hipFunction_t fn = nullptr;
  int resident = 2;  // workgroups per CU
};

struct JitCache {
  std::mutex mu;
  std::unordered_map<std::string, Compiled> map;  // key: device + source
  std::atomic<int64_t> compiled{0}, hits{0};
  std::atomic<int64_t> compile_us{0};
};

// Held for the life of the process (never destroyed): a background compile
// may still be using it while static objects are torn down at exit.
JitCache &cache() {
  static JitCache *c = new JitCache;
  return *c;
}

// Directory of libqmhip.so, where the build puts qm_jitc beside it.
std::string helper_path() {
  Dl_info info{};
  if (!dladdr((const void *)&helper_path, &info) || !info.dli_fname) return {};
  std::string lib = info.dli_fname;
  const size_t slash = lib.rfind('/');
  return (slash == std::string::npos ? std::string(".") : lib.substr(0, slash)) + "/qm_jitc";
}

std::vector<char> read_file(const std::string &path) {
  std::vector<char> out;
  FILE *f = std::fopen(path.c_str(), "rb");
  if (!f) return out;
  char buf[65536];
  size_t n;
  while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) out.insert(out.end(), buf, buf + n);
  std::fclose(f);
  return out;
}

// Compiles the sources in up to `workers` qm_jitc child processes (the
// sources and code objects pass through a private temporary directory).
// Returns false, with nothing compiled, when the helper cannot be started.
bool compile_spawned(const std::vector<std::string> &srcs, size_t workers,
                     std::vector<std::vector<char>> &codes, std::vector<std::string> &errs) {
  const std::string helper = helper_path();
  if (helper.empty() || access(helper.c_str(), X_OK) != 0) return false;
  const char *tmp = std::getenv("TMPDIR");
  std::string templ = std::string(tmp && *tmp ? tmp : "/tmp") + "/qmjit.XXXXXX";
  std::vector<char> dir_buf(templ.begin(), templ.end());
  dir_buf.push_back('\0');
  if (!mkdtemp(dir_buf.data())) return false;
  const std::string dir = dir_buf.data();
  auto file = [&](size_t k, const char *ext) { return dir + "/k" + std::to_string(k) + ext; };
  bool ok = true;
  for (size_t k = 0; k < srcs.size() && ok; k++) {
    FILE *f = std::fopen(file(k, ".hip").c_str(), "wb");
    ok = f && std::fwrite(srcs[k].data(), 1, srcs[k].size(), f) == srcs[k].size();
    if (f) std::fclose(f);
  }
  std::vector<pid_t> pids;
  for (size_t w = 0; w < workers && ok; w++) {
    std::vector<std::string> args{helper, dir};
    for (size_t k = w; k < srcs.size(); k += workers) args.push_back(std::to_string(k));
    std::vector<char *> argv;
    for (std::string &a : args) argv.push_back(&a[0]);
    argv.push_back(nullptr);
    pid_t pid = 0;
    if (posix_spawn(&pid, helper.c_str(), nullptr, nullptr, argv.data(), environ) != 0) {
      ok = false;
      break;
    }
    pids.push_back(pid);
  }
  for (pid_t pid : pids) {
    int st = 0;
    while (waitpid(pid, &st, 0) < 0 && errno == EINTR) {
    }
  }
  const bool spawned = ok;
  for (size_t k = 0; k < srcs.size(); k++) {
    if (spawned) {
      codes[k] = read_file(file(k, ".co"));
      if (codes[k].empty()) {
        const std::vector<char> e = read_file(file(k, ".err"));
        errs[k] = e.empty() ? "the compile helper produced no code object" : std::string(e.begin(), e.end());
      }
    }
    for (const char *ext : {".hip", ".co", ".err", ".tmp"}) std::remove(file(k, ext).c_str());
  }
  rmdir(dir.c_str());
  return spawned;
}

// Compiles every source: in helper processes when there are several (hipRTC
// compiles one program at a time per process), else on threads in-process.
// QM_JIT_PROCS caps the helper count (0: compile in-process).
std::vector<std::vector<char>> compile_all(const std::vector<std::string> &srcs) {
  std::vector<std::vector<char>> codes(srcs.size());
  std::vector<std::string> errs(srcs.size());
  size_t procs = std::min<size_t>(std::max(1u, std::thread::hardware_concurrency()), 16);
  if (const char *e = std::getenv("QM_JIT_PROCS")) procs = (size_t)std::max(0, std::atoi(e));
  const size_t workers = std::min(srcs.size(), procs);
  if (workers < 2 || !compile_spawned(srcs, workers, codes, errs)) {
    std::atomic<size_t> next{0};
    std::vector<std::thread> th;
    for (size_t w = 0; w < std::max<size_t>(std::min<size_t>(srcs.size(), 16), 1); w++)
      th.emplace_back([&] {
        for (size_t k; (k = next++) < srcs.size();) codes[k] = rtc_compile(srcs[k], errs[k]);
      });
    for (auto &x : th) x.join();
  }
  for (size_t k = 0; k < srcs.size(); k++)
    if (codes[k].empty()) fail(QM_ERR_DEVICE, "pass compilation failed: " + errs[k]);
  return codes;
}

}  // namespace

int jit_default_variant(int precision) { return jit_variant(precision); }

bool jit_supported(const FusedPlan &plan) {
  for (const FusedPass &p : plan.passes)
    for (const TileOp &op : p.ops)
      if (op.kind >= G_PHASEPROD) return false;
  return true;
}

// Sources (cache keys) and coefficients of a plan's pass kernels.
static void plan_keys(const FusedPlan &plan, int precision, int device, JitProgram &jp,
                      std::vector<std::string> &keys) {
  keys.resize(plan.passes.size());
  for (size_t i = 0; i < plan.passes.size(); i++) {
    // Coefficient indices in the source are relative to the pass (the source,
    // and so the cache key, does not depend on where the pass sits).
    std::vector<double> local;
    std::vector<int32_t> *prov = nullptr;
    if (plan.shape.state_bits > 0) {
      jp.prov.resize(plan.passes.size());
      prov = &jp.prov[i];
    }
    std::string src = pass_source(plan.passes[i], plan.shape, precision, local, prov);
    jp.coef_offset.push_back((int64_t)jp.coef.size());
    jp.coef.insert(jp.coef.end(), local.begin(), local.end());
    keys[i] = std::to_string(device) + "\n" + kJitPrelude + src;
  }
}

bool jit_lookup(const FusedPlan &plan, int precision, int device, JitProgram *out) {
  check_device_variant(variant_of(plan.shape, precision));
  JitProgram jp;
  std::vector<std::string> keys;
  plan_keys(plan, precision, device, jp, keys);
  JitCache &c = cache();
  std::lock_guard<std::mutex> lk(c.mu);
  for (const std::string &k : keys)
    if (!c.map.count(k)) return false;
  for (const std::string &k : keys) {
    jp.fns.push_back((void *)c.map[k].fn);
    jp.resident.push_back(c.map[k].resident);
  }
  c.hits += (int64_t)keys.size();
  *out = std::move(jp);
  return true;
}

namespace {
// Background compilation of plans whose kernels a later execution will find
// in the cache: one worker at a time, plans already queued not queued again.
struct AsyncJit {
  std::mutex mu;
  std::condition_variable idle;
  int running = 0;
  std::unordered_map<std::string, int> queued;  // by the plan's first pass key + pass count
};
AsyncJit &async_jit() {
  static AsyncJit *a = new AsyncJit;
  return *a;
}
std::mutex &async_serial() {
  static std::mutex *m = new std::mutex;
  return *m;
}
}  // namespace

// Background job: `make` produces the plan (possibly searching for it) and the
// plan's missing kernels are then compiled; jobs of one tag run once at a time,
// and at most four jobs are in flight.
static void jit_job(const std::string &tag, std::function<FusedPlan()> make, int precision, int device) {
  AsyncJit &a = async_jit();
  {
    std::lock_guard<std::mutex> lk(a.mu);
    if (a.queued.count(tag) || a.running >= 4) return;  // already queued, or the queue is full
    a.queued[tag] = 1;
    a.running++;
  }
  std::thread([make, precision, device, tag] {
    try {
      const FusedPlan plan = make();
      if (!plan.passes.empty() && jit_supported(plan)) {
        std::lock_guard<std::mutex> serial(async_serial());
        (void)hipSetDevice(device);
        (void)jit_prepare(plan, precision, device);
      }
    } catch (...) {
      // A failed background plan or compile only means the next execution
      // plans or compiles again (and reports the failure) or runs interpreted.
    }
    AsyncJit &a = async_jit();
    std::lock_guard<std::mutex> lk(a.mu);
    a.queued.erase(tag);
    a.running--;
    a.idle.notify_all();
  }).detach();
}

void jit_prepare_async(const FusedPlan &plan, int precision, int device) {
  if (plan.passes.empty()) return;
  std::string tag;
  {
    JitProgram jp;
    std::vector<std::string> keys;
    plan_keys(plan, precision, device, jp, keys);
    for (const std::string &k : keys) tag += std::to_string(std::hash<std::string>{}(k)) + ",";
  }
  jit_job(tag, [plan] { return plan; }, precision, device);
}

void jit_plan_async(const std::string &structure, std::function<FusedPlan()> make, int precision,
                    int device) {
  jit_job("plan:" + std::to_string(device) + ":" + std::to_string(std::hash<std::string>{}(structure)),
          std::move(make), precision, device);
}

void jit_drain() {
  AsyncJit &a = async_jit();
  std::unique_lock<std::mutex> lk(a.mu);
  a.idle.wait(lk, [&] { return a.running == 0; });
}

JitProgram jit_prepare(const FusedPlan &plan, int precision, int device) {
  check_device_variant(variant_of(plan.shape, precision));
  JitProgram jp;
  std::vector<std::string> keys;
  plan_keys(plan, precision, device, jp, keys);
  JitCache &c = cache();
  std::vector<size_t> missing;
  {
    std::lock_guard<std::mutex> lk(c.mu);
    std::unordered_map<std::string, int> seen;
    for (size_t i = 0; i < keys.size(); i++) {
      if (c.map.count(keys[i])) {
        c.hits++;
      } else if (!seen.count(keys[i])) {
        seen[keys[i]] = 1;
        missing.push_back(i);
      }
    }
  }
  if (!missing.empty()) {
    const auto t0 = std::chrono::steady_clock::now();
    // Highest occupancy (3, then 2 workgroups per CU, at most what the
    // exchange buffer leaves room for in the CU's LDS) whose register
    // allocation does not spill, per pass.
    std::vector<size_t> todo(missing.size());
    for (size_t k = 0; k < todo.size(); k++) todo[k] = k;
    std::vector<Compiled> done(missing.size());
    const bool prefetch = (variant_of(plan.shape, precision) & 64) != 0;
    static const int fixed_occ = [] {
      const char *e = std::getenv("QM_JIT_OCC");  // tuning knob: compile for this occupancy only
      return e ? std::atoi(e) : 0;
    }();
    int lds_per_cu = 0;
    QM_HIP(hipDeviceGetAttribute(&lds_per_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, device));
    const int lds_occ =
        std::max(1, (int)((size_t)lds_per_cu / std::max<size_t>(jit_lds_bytes(precision, plan.shape), 1)));
    const int occ_hi = fixed_occ > 0 ? fixed_occ : std::min(prefetch ? 1 : 3, lds_occ);
    const int occ_lo = fixed_occ > 0 ? fixed_occ : std::min(prefetch ? 1 : 2, lds_occ);
    for (int occ = occ_hi; occ >= occ_lo && !todo.empty(); occ--) {
      std::vector<std::string> srcs;
      for (size_t k : todo) {
        const std::string &key = keys[missing[k]];
        srcs.push_back("#define QM_OCC " + std::to_string(occ) + "\n" + key.substr(key.find('\n') + 1));
      }
      std::vector<std::vector<char>> codes = compile_all(srcs);
      std::vector<size_t> again;
      for (size_t i = 0; i < todo.size(); i++) {
        Compiled cp;
        QM_HIP(hipModuleLoadData(&cp.mod, codes[i].data()));
        QM_HIP(hipModuleGetFunction(&cp.fn, cp.mod, "qm_pass"));
        int scratch = 0;
        QM_HIP(hipFuncGetAttribute(&scratch, HIP_FUNC_ATTRIBUTE_LOCAL_SIZE_BYTES, cp.fn));
        if (scratch > 0 && occ > occ_lo) {
          (void)hipModuleUnload(cp.mod);
          again.push_back(todo[i]);
          continue;
        }
        int per_cu = 0;
        QM_HIP(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, cp.fn, 1 << plan.shape.L, jit_lds_bytes(precision, plan.shape)));
        cp.resident = std::max(per_cu, 1);
        done[todo[i]] = cp;
      }
      todo.swap(again);
    }
    std::lock_guard<std::mutex> lk(c.mu);
    for (size_t k = 0; k < missing.size(); k++) {
      c.map[keys[missing[k]]] = done[k];
      c.compiled++;
    }
    c.compile_us += (int64_t)std::chrono::duration_cast<std::chrono::microseconds>(
                        std::chrono::steady_clock::now() - t0)
                        .count();
  }
  std::lock_guard<std::mutex> lk(c.mu);
  for (const std::string &k : keys) {
    jp.fns.push_back((void *)c.map[k].fn);
    jp.resident.push_back(c.map[k].resident);
  }
  return jp;
}

void jit_launch_batch(int compute_units, int precision, const FusedShape &shape, const JitProgram &jp,
                      size_t pass, const double *d_coef, int64_t items, void *states, hipStream_t st) {
  const uint64_t tiles = (uint64_t)items << (shape.state_bits - shape.K);
  const size_t lds = jit_lds_bytes(precision, shape);
  const uint64_t want = (uint64_t)compute_units * (uint64_t)jp.resident[pass];
  const unsigned blocks = (unsigned)std::min<uint64_t>(tiles, std::max<uint64_t>(want, 1));
  unsigned long long nt = tiles;
  void *args[] = {&states, &d_coef, &nt};
  QM_HIP(hipModuleLaunchKernel((hipFunction_t)jp.fns[pass], blocks, 1, 1, 1u << shape.L, 1, 1, (unsigned)lds, st,
                               args, nullptr));
}

void jit_launch(State *s, const FusedShape &shape, const JitProgram &jp, size_t pass,
                const double *d_coef, hipStream_t st, const void *src, void *dst) {
  const uint64_t tiles = 1ull << (s->qubits - shape.K);
  const size_t lds = jit_lds_bytes(s->precision, shape);
  const uint64_t want = (uint64_t)s->ctx->compute_units * (uint64_t)jp.resident[pass];
  const unsigned blocks = (unsigned)std::min<uint64_t>(tiles, std::max<uint64_t>(want, 1));
  const double *cf = d_coef + jp.coef_offset[pass];
  unsigned long long nt = tiles;
  void *args_in[] = {&dst, &cf, &nt};
  void *args_out[] = {&src, &dst, &cf, &nt};
  QM_HIP(hipModuleLaunchKernel((hipFunction_t)jp.fns[pass], blocks, 1, 1, 1u << shape.L, 1, 1,
                               (unsigned)lds, st, src == dst ? args_in : args_out, nullptr));
}

int jit_compile_check(const FusedPlan &plan, int precision) {
  const char *dump = std::getenv("QM_JIT_DUMP");  // debugging aid: directory for the sources
  std::vector<std::string> srcs;
  int index = 0;
  for (const FusedPass &p : plan.passes) {
    std::vector<double> coef;
    const std::string src = pass_source(p, plan.shape, precision, coef);
    if (dump) {
      FILE *f = std::fopen((std::string(dump) + "/pass" + std::to_string(index) + ".hip").c_str(), "w");
      if (f) {
        std::fputs(src.c_str(), f);
        std::fclose(f);
      }
      f = std::fopen((std::string(dump) + "/pass" + std::to_string(index) + ".coef").c_str(), "w");
      if (f) {
        for (double v : coef) std::fprintf(f, "%.17g\n", v);
        std::fclose(f);
      }
      index++;
    }
    srcs.push_back("#define QM_OCC 2\n" + std::string(kJitPrelude) + src);
  }
  compile_all(srcs);
  return (int)plan.passes.size();
}

void jit_stats(int64_t *compiled, int64_t *hits, double *seconds) {
  JitCache &c = cache();
  if (compiled) *compiled = c.compiled.load();
  if (hits) *hits = c.hits.load();
  if (seconds) *seconds = (double)c.compile_us.load() * 1e-6;
}

}  // namespace qm
