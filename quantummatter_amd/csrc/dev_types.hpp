// Device amplitude types shared by the kernel translation units.
#pragma once

#include <cstdint>

namespace qm {

struct alignas(16) C128 {
  double re, im;
};
struct alignas(8) C64 {
  float re, im;
};
template <typename T> struct Real;
template <> struct Real<C128> { using type = double; };
template <> struct Real<C64> { using type = float; };

constexpr int kBlock = 256;

}  // namespace qm
