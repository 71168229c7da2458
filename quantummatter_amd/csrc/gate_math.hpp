// Per-pair gate arithmetic shared by the HIP kernels and the host reference.
//
// Each form is the reference's update (a*x + b*y, c*x + d*y) with
// System.Numerics.Complex multiplication, re = ar*xr - ai*xi,
// im = ai*xr + ar*xi (CpuStateVectorBackend.cs:113-116), reduced for the
// coefficient pattern of the gate kind.  Every reduction drops only products
// with an exact zero coefficient, so in double precision the result equals
// the general form bit for bit (up to the sign of a zero).  The library is
// compiled with -ffp-contract=off: no product is fused into an add.
#pragma once

#if defined(__HIPCC__)
#define QM_HD __host__ __device__ __forceinline__
#else
#define QM_HD inline
#endif

namespace qm {

template <typename F>
QM_HD void pair_mat(const F *c, F &xr, F &xi, F &yr, F &yi) {
  const F x0 = xr, x1 = xi, y0 = yr, y1 = yi;
  xr = (c[0] * x0 - c[1] * x1) + (c[2] * y0 - c[3] * y1);
  xi = (c[1] * x0 + c[0] * x1) + (c[3] * y0 + c[2] * y1);
  yr = (c[4] * x0 - c[5] * x1) + (c[6] * y0 - c[7] * y1);
  yi = (c[5] * x0 + c[4] * x1) + (c[7] * y0 + c[6] * y1);
}

// Rx: a = d = (c, 0), b = c' = (-0, -s).
template <typename F>
QM_HD void pair_rx(F c, F s, F &xr, F &xi, F &yr, F &yi) {
  const F x0 = xr, x1 = xi, y0 = yr, y1 = yi;
  xr = c * x0 + s * y1;
  xi = c * x1 - s * y0;
  yr = s * x1 + c * y0;
  yi = c * y1 - s * x0;
}

// Ry: a = d = (c, 0), b = (-s, 0), c' = (s, 0).
template <typename F>
QM_HD void pair_ry(F c, F s, F &xr, F &xi, F &yr, F &yi) {
  const F x0 = xr, x1 = xi, y0 = yr, y1 = yi;
  xr = c * x0 - s * y0;
  xi = c * x1 - s * y1;
  yr = s * x0 + c * y0;
  yi = s * x1 + c * y1;
}

// H: a = b = c' = (h, 0), d = (-h, 0).
template <typename F>
QM_HD void pair_h(F h, F &xr, F &xi, F &yr, F &yi) {
  const F x0 = xr, x1 = xi, y0 = yr, y1 = yi;
  xr = h * x0 + h * y0;
  xi = h * x1 + h * y1;
  yr = h * x0 - h * y0;
  yi = h * x1 - h * y1;
}

// Y: b = (-0, -1), c' = (0, 1): an exact signed permutation.
template <typename F>
QM_HD void pair_y(F &xr, F &xi, F &yr, F &yi) {
  const F x0 = xr, x1 = xi, y0 = yr, y1 = yi;
  xr = y1;
  xi = -y0;
  yr = -x1;
  yi = x0;
}

// Diagonal phase p applied to one amplitude (state[i] *= p).
template <typename F>
QM_HD void mul_phase(F pr, F pi, F &xr, F &xi) {
  const F x0 = xr, x1 = xi;
  xr = x0 * pr - x1 * pi;
  xi = x1 * pr + x0 * pi;
}

// ---------------------------------------------------------------------------
// Fused-multiply-add forms used by the fused tile passes.  Same products and
// sums as above with one rounding less per fused pair, so the results differ
// from the reference's rounding by at most an ulp per update (far inside the
// 1e-12 amplitude tolerance) and cost a third fewer instructions.
// ---------------------------------------------------------------------------
template <typename F>
QM_HD F fma_(F a, F b, F c) {
  return __builtin_fma(a, b, c);
}
template <>
QM_HD float fma_<float>(float a, float b, float c) {
  return __builtin_fmaf(a, b, c);
}

template <typename F>
QM_HD void pair_mat_fma(const F *c, F &xr, F &xi, F &yr, F &yi) {
  const F x0 = xr, x1 = xi, y0 = yr, y1 = yi;
  xr = fma_(-c[3], y1, fma_(c[2], y0, fma_(-c[1], x1, c[0] * x0)));
  xi = fma_(c[2], y1, fma_(c[3], y0, fma_(c[0], x1, c[1] * x0)));
  yr = fma_(-c[7], y1, fma_(c[6], y0, fma_(-c[5], x1, c[4] * x0)));
  yi = fma_(c[6], y1, fma_(c[7], y0, fma_(c[4], x1, c[5] * x0)));
}

template <typename F>
QM_HD void pair_rx_fma(F c, F s, F &xr, F &xi, F &yr, F &yi) {
  const F x0 = xr, x1 = xi, y0 = yr, y1 = yi;
  xr = fma_(s, y1, c * x0);
  xi = fma_(-s, y0, c * x1);
  yr = fma_(s, x1, c * y0);
  yi = fma_(-s, x0, c * y1);
}

template <typename F>
QM_HD void pair_ry_fma(F c, F s, F &xr, F &xi, F &yr, F &yi) {
  const F x0 = xr, x1 = xi, y0 = yr, y1 = yi;
  xr = fma_(-s, y0, c * x0);
  xi = fma_(-s, y1, c * x1);
  yr = fma_(s, x0, c * y0);
  yi = fma_(s, x1, c * y1);
}

template <typename F>
QM_HD void pair_h_fma(F h, F &xr, F &xi, F &yr, F &yi) {
  const F x0 = xr, x1 = xi, y0 = yr, y1 = yi;
  xr = fma_(h, y0, h * x0);
  xi = fma_(h, y1, h * x1);
  yr = fma_(-h, y0, h * x0);
  yi = fma_(-h, y1, h * x1);
}

template <typename F>
QM_HD void mul_phase_fma(F pr, F pi, F &xr, F &xi) {
  const F x0 = xr, x1 = xi;
  xr = fma_(-x1, pi, x0 * pr);
  xi = fma_(x0, pi, x1 * pr);
}

}  // namespace qm
