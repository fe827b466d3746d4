// One MT19937 block step (twist, then temper all 624 words), written as plain
// loops the compiler vectorises: the first 227 updates read only old state,
// the next 396 read words updated at least 227 positions earlier, the last one
// reads the new x[0].  On x86-64 an AVX2 build of the same code is chosen at
// run time when the CPU supports it.
#include "mt19937.hpp"

namespace rsa_amd {
namespace {

__attribute__((always_inline)) inline void block_body(uint32_t* __restrict__ x, uint32_t* __restrict__ out) {
  constexpr int N = MT19937::N, M = MT19937::M;
  constexpr uint32_t A = MT19937::MATRIX_A, UP = MT19937::UPPER, LO = MT19937::LOWER;
  for (int k = 0; k < N - M; k++) {
    const uint32_t y = (x[k] & UP) | (x[k + 1] & LO);
    x[k] = x[k + M] ^ (y >> 1) ^ ((0u - (y & 1u)) & A);
  }
  for (int k = N - M; k < N - 1; k++) {
    const uint32_t y = (x[k] & UP) | (x[k + 1] & LO);
    x[k] = x[k + (M - N)] ^ (y >> 1) ^ ((0u - (y & 1u)) & A);
  }
  const uint32_t y = (x[N - 1] & UP) | (x[0] & LO);
  x[N - 1] = x[M - 1] ^ (y >> 1) ^ ((0u - (y & 1u)) & A);
  for (int i = 0; i < N; i++) out[i] = MT19937::temper(x[i]);
}

void block_generic(uint32_t* x, uint32_t* out) { block_body(x, out); }

#if defined(__x86_64__)
__attribute__((target("avx2"))) void block_avx2(uint32_t* x, uint32_t* out) { block_body(x, out); }
#endif

}  // namespace

void mt_block(uint32_t* x, uint32_t* out) {
#if defined(__x86_64__)
  static const bool avx2 = __builtin_cpu_supports("avx2");
  if (avx2) {
    block_avx2(x, out);
    return;
  }
#endif
  block_generic(x, out);
}

}  // namespace rsa_amd
