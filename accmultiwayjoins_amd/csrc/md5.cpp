// MD5 (RFC 1321, section 3): 512-bit blocks, four rounds of 16 steps, little-endian words.
// The sine table T[i] = floor(2^32 * |sin(i + 1)|) is computed at start-up as the RFC defines it.
#include "md5.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace amj {

namespace {

struct Tables {
  uint32_t t[64];
  Tables() {
    for (int i = 0; i < 64; ++i) t[i] = (uint32_t)std::floor(std::fabs(std::sin((double)(i + 1))) * 4294967296.0);
  }
};
const Tables kT;

// per-round left-rotation amounts (RFC 1321 3.4)
const int kS[4][4] = {{7, 12, 17, 22}, {5, 9, 14, 20}, {4, 11, 16, 23}, {6, 10, 15, 21}};

inline uint32_t rotl(uint32_t x, int c) { return (x << c) | (x >> (32 - c)); }

}  // namespace

Md5::Md5() {
  h_[0] = 0x67452301u;
  h_[1] = 0xefcdab89u;
  h_[2] = 0x98badcfeu;
  h_[3] = 0x10325476u;
}

void Md5::block(const uint8_t* p) {
  uint32_t x[16];
  for (int i = 0; i < 16; ++i)
    x[i] = (uint32_t)p[4 * i] | ((uint32_t)p[4 * i + 1] << 8) | ((uint32_t)p[4 * i + 2] << 16) |
           ((uint32_t)p[4 * i + 3] << 24);
  uint32_t a = h_[0], b = h_[1], c = h_[2], d = h_[3];
  for (int i = 0; i < 64; ++i) {
    const int r = i / 16;
    uint32_t f;
    int k;
    switch (r) {
      case 0: f = (b & c) | (~b & d); k = i; break;
      case 1: f = (b & d) | (c & ~d); k = (5 * i + 1) & 15; break;
      case 2: f = b ^ c ^ d; k = (3 * i + 5) & 15; break;
      default: f = c ^ (b | ~d); k = (7 * i) & 15; break;
    }
    const uint32_t nb = b + rotl(a + f + x[k] + kT.t[i], kS[r][i & 3]);
    a = d;
    d = c;
    c = b;
    b = nb;
  }
  h_[0] += a;
  h_[1] += b;
  h_[2] += c;
  h_[3] += d;
}

void Md5::update(const void* data, size_t n) {
  const uint8_t* p = (const uint8_t*)data;
  len_ += n;
  while (n > 0) {
    const size_t take = std::min<size_t>(64 - fill_, n);
    std::memcpy(buf_ + fill_, p, take);
    fill_ += take;
    p += take;
    n -= take;
    if (fill_ == 64) {
      block(buf_);
      fill_ = 0;
    }
  }
}

std::string Md5::hexdigest() {
  const uint64_t bits = len_ * 8;
  const uint8_t one = 0x80, zero = 0;
  update(&one, 1);
  while (fill_ != 56) update(&zero, 1);
  uint8_t l[8];
  for (int i = 0; i < 8; ++i) l[i] = (uint8_t)(bits >> (8 * i));
  update(l, 8);
  static const char* hex = "0123456789abcdef";
  std::string s;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      const uint8_t byte = (uint8_t)(h_[i] >> (8 * j));
      s.push_back(hex[byte >> 4]);
      s.push_back(hex[byte & 15]);
    }
  return s;
}

}  // namespace amj
