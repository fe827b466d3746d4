// MT19937 word stream of the reference's RNG (rsa_random.hxx:37-42:
// std::mt19937 seeded with std::seed_seq{rank, seed}), written from the
// published definitions: Matsumoto & Nishimura (1998) for the generator and
// [rand.util.seedseq] / [rand.eng.mers] of ISO C++ for seed_seq::generate and
// the engine's seeding.  The stream is produced in whole 624-word blocks.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

namespace rsa_amd {

// twist state x[624] once and write the 624 tempered words to out (mt19937.cpp)
void mt_block(uint32_t* x, uint32_t* out);

struct MT19937 {
  static constexpr int N = 624, M = 397;
  static constexpr uint32_t MATRIX_A = 0x9908b0dfu, UPPER = 0x80000000u, LOWER = 0x7fffffffu;
  uint32_t x[N];

  // seed_seq{v0, v1}.generate(x, x + 624) followed by mersenne_twister_engine::seed(seq)
  void seed_seq2(uint32_t v0, uint32_t v1) {
    const uint32_t v[2] = {v0, v1};
    const size_t n = N, s = 2;
    for (size_t i = 0; i < n; i++) x[i] = 0x8b8b8b8bu;
    const size_t t = (n >= 623) ? 11 : (n >= 68) ? 7 : (n >= 39) ? 5 : (n >= 7) ? 3 : (n - 1) / 2;
    const size_t p = (n - t) / 2, q = p + t;
    const size_t m = (s + 1 > n) ? s + 1 : n;
    auto T = [](uint32_t a) { return a ^ (a >> 27); };
    for (size_t k = 0; k < m; k++) {
      const uint32_t r1 = 1664525u * T(x[k % n] ^ x[(k + p) % n] ^ x[(k + n - 1) % n]);
      uint32_t r2 = r1;
      if (k == 0)
        r2 += (uint32_t)s;
      else if (k <= s)
        r2 += (uint32_t)(k % n) + v[k - 1];
      else
        r2 += (uint32_t)(k % n);
      x[(k + p) % n] += r1;
      x[(k + q) % n] += r2;
      x[k % n] = r2;
    }
    for (size_t k = m; k < m + n; k++) {
      const uint32_t r3 = 1566083941u * T(x[k % n] + x[(k + p) % n] + x[(k + n - 1) % n]);
      const uint32_t r4 = r3 - (uint32_t)(k % n);
      x[(k + p) % n] ^= r3;
      x[(k + q) % n] ^= r4;
      x[k % n] = r4;
    }
    // engine seeding from the generated words: an all-zero significant state
    // (top bit of x[0] and x[1..623]) is replaced by 2^31
    bool zero = (x[0] & UPPER) == 0;
    for (int i = 1; i < N && zero; i++)
      if (x[i] != 0) zero = false;
    if (zero) x[0] = UPPER;
  }

  // the classic integer seeding (used only for the default-seed known answer)
  void seed_u32(uint32_t s) {
    x[0] = s;
    for (int i = 1; i < N; i++) x[i] = 1812433253u * (x[i - 1] ^ (x[i - 1] >> 30)) + (uint32_t)i;
  }

  static inline uint32_t temper(uint32_t z) {
    z ^= (z >> 11);
    z ^= (z << 7) & 0x9d2c5680u;
    z ^= (z << 15) & 0xefc60000u;
    z ^= (z >> 18);
    return z;
  }

  // next 624 output words (the engine twists before its first output)
  void next_block(uint32_t* out) { mt_block(x, out); }
};

// A growing window of the stream: words [base, base + size) are kept.
struct MTStream {
  MT19937 mt;
  std::vector<uint32_t> buf;  // buf[i] = word base + i
  uint64_t base = 0;

  void reset(uint64_t seed, uint64_t rank = 0) {
    mt.seed_seq2((uint32_t)rank, (uint32_t)seed);
    buf.clear();
    base = 0;
  }
  uint64_t end() const { return base + buf.size(); }
  // make sure words [base, upto) exist
  void ensure(uint64_t upto) {
    while (end() < upto) {
      const size_t old = buf.size();
      buf.resize(old + MT19937::N);
      mt.next_block(buf.data() + old);
    }
  }
  uint32_t at(uint64_t pos) {
    ensure(pos + 1);
    return buf[pos - base];
  }
  // forget words below pos (keeps whole blocks only)
  void drop_below(uint64_t pos) {
    if (pos <= base) return;
    const uint64_t k = ((pos - base) / MT19937::N) * MT19937::N;
    if (k == 0) return;
    buf.erase(buf.begin(), buf.begin() + (ptrdiff_t)k);
    base += k;
  }
};

// URNG adaptor over MTStream starting at a position; counts the words used.
// min()/max()/result_type are those of std::mt19937 so libstdc++
// distributions take the same code path as in the reference.
struct StreamURNG {
  using result_type = unsigned long;  // std::mt19937::result_type (uint_fast32_t) on LP64
  MTStream* s;
  uint64_t pos;
  static constexpr result_type min() { return 0; }
  static constexpr result_type max() { return 0xffffffffUL; }
  result_type operator()() { return s->at(pos++); }
};

}  // namespace rsa_amd
