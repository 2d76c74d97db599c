// SIMD-oriented Fast Mersenne Twister, period 2^19937 - 1 (Saito and
// Matsumoto, "SIMD-oriented Fast Mersenne Twister: a 128-bit Pseudorandom
// Number Generator", MCQMC 2006), in portable scalar form: 156 128-bit words,
// the published recursion and parameters of that period (pick-up position
// 122, shifts 18 / 1 byte / 11 / 1 byte, masks and parity words below).
//
// The reference draws ParticleSystem::randPos (particleSystem.h:115-120) from
// this generator: init_gen_rand(0) when the reaction tables are set up
// (particleSystem.cpp:342, 437, 536), then unirandOOF(-1, 1) per coordinate
// (stats.h:35-37: (float)genrand_real3() * (hi - lo) + lo, genrand_real3 on
// the open interval (0, 1)).
#include <cstdint>

#include "smd_setup.h"

namespace smd {

namespace {

constexpr int kPos1 = 122;
constexpr int kSl1 = 18, kSl2Bytes = 1, kSr1 = 11, kSr2Bytes = 1;
constexpr uint32_t kMsk[4] = {0xdfffffefu, 0xddfecb7fu, 0xbffaffffu, 0xbffffff6u};
constexpr uint32_t kParity[4] = {0x00000001u, 0x00000000u, 0x00000000u, 0x13c9e684u};

// 128-bit value shifted left / right by whole bytes (word 0 least significant)
void shl128(uint32_t out[4], const uint32_t in[4], int bytes) {
  const uint64_t hi = ((uint64_t)in[3] << 32) | in[2];
  const uint64_t lo = ((uint64_t)in[1] << 32) | in[0];
  const int s = bytes * 8;
  const uint64_t oh = (hi << s) | (lo >> (64 - s));
  const uint64_t ol = lo << s;
  out[0] = (uint32_t)ol;
  out[1] = (uint32_t)(ol >> 32);
  out[2] = (uint32_t)oh;
  out[3] = (uint32_t)(oh >> 32);
}

void shr128(uint32_t out[4], const uint32_t in[4], int bytes) {
  const uint64_t hi = ((uint64_t)in[3] << 32) | in[2];
  const uint64_t lo = ((uint64_t)in[1] << 32) | in[0];
  const int s = bytes * 8;
  const uint64_t ol = (lo >> s) | (hi << (64 - s));
  const uint64_t oh = hi >> s;
  out[0] = (uint32_t)ol;
  out[1] = (uint32_t)(ol >> 32);
  out[2] = (uint32_t)oh;
  out[3] = (uint32_t)(oh >> 32);
}

}  // namespace

void Sfmt::seed(uint32_t s) {
  uint32_t* w = &st[0][0];
  w[0] = s;
  for (int i = 1; i < kWords32; ++i) w[i] = 1812433253u * (w[i - 1] ^ (w[i - 1] >> 30)) + (uint32_t)i;
  idx = kWords32;
  // period certification: the state must not lie in the sub-space the
  // parity words exclude; flip the lowest parity bit if it does
  uint32_t inner = 0;
  for (int i = 0; i < 4; ++i) inner ^= w[i] & kParity[i];
  for (int i = 16; i > 0; i >>= 1) inner ^= inner >> i;
  if (inner & 1u) return;
  for (int i = 0; i < 4; ++i) {
    for (int b = 0; b < 32; ++b) {
      const uint32_t bit = 1u << b;
      if (bit & kParity[i]) {
        w[i] ^= bit;
        return;
      }
    }
  }
}

void Sfmt::refill() {
  // w[i] = a ^ (a << 8 SL2 bits) ^ ((b >> SR1) & MSK) ^ (c >> 8 SR2 bits) ^ (d << SL1)
  // with a = w[i], b = w[i + POS1], c = w[i - 2], d = w[i - 1] (indices mod N)
  int c = kWords128 - 2, d = kWords128 - 1;
  for (int i = 0; i < kWords128; ++i) {
    const int b = (i + kPos1) % kWords128;
    uint32_t x[4], y[4], r[4];
    shl128(x, st[i], kSl2Bytes);
    shr128(y, st[c], kSr2Bytes);
    for (int k = 0; k < 4; ++k)
      r[k] = st[i][k] ^ x[k] ^ ((st[b][k] >> kSr1) & kMsk[k]) ^ y[k] ^ (st[d][k] << kSl1);
    for (int k = 0; k < 4; ++k) st[i][k] = r[k];
    c = d;
    d = i;
  }
  idx = 0;
}

uint32_t Sfmt::next32() {
  if (idx >= kWords32) refill();
  const uint32_t v = (&st[0][0])[idx];
  ++idx;
  return v;
}

double Sfmt::real_open() { return ((double)next32() + 0.5) * (1.0 / 4294967296.0); }

float Sfmt::uniform_open(float lo, float hi) {
  const float u = (float)real_open();
  const float span = hi - lo;
  const float scaled = u * span;  // separate rounding of the product and the sum
  return scaled + lo;
}

}  // namespace smd
