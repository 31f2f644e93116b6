// MD5 message digest (RFC 1321), used for the reference's result fingerprint.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>

namespace amj {

class Md5 {
 public:
  Md5();
  void update(const void* data, size_t n);
  // 32 lowercase hex digits; the object must not be updated afterwards
  std::string hexdigest();

 private:
  void block(const uint8_t* p);
  uint32_t h_[4];
  uint8_t buf_[64];
  uint64_t len_ = 0;  // bytes
  size_t fill_ = 0;
};

}  // namespace amj
