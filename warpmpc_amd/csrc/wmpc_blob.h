// One device allocation holding a plan's read-only tables (host side).
#pragma once

#include <cstring>
#include <string>
#include <vector>

#include "wmpc_common.h"

namespace wmpc {

// Tables are appended as 32-bit words, each starting on a 256-byte boundary
// and followed by one zero word (a zero-length table still has a valid
// address).  upload() copies the blob to the current device; at(i) is then the
// device address of the i-th table.
class Blob {
 public:
  template <class T>
  size_t add(const T* src, size_t count) {
    static_assert(sizeof(T) == 4 || sizeof(T) == 8 || sizeof(T) == 16, "32-bit element types only");
    offs_.push_back(words_.size());
    const size_t w = count * sizeof(T) / 4;
    const size_t at = words_.size();
    words_.resize(at + w);
    if (w) std::memcpy(words_.data() + at, src, w * 4);
    words_.push_back(0);
    while (words_.size() % 64) words_.push_back(0);
    return offs_.size() - 1;
  }
  template <class T>
  size_t add(const std::vector<T>& v) {
    return add(v.data(), v.size());
  }
  size_t add_f32(const double* src, size_t count) {
    std::vector<float> v(count);
    for (size_t i = 0; i < count; ++i) v[i] = (float)src[i];
    return add(v);
  }
  // Allocates and fills the device copy; on failure sets the error text.
  int upload(void** block, const char* what) {
    *block = nullptr;
    hipError_t e = hipMalloc(block, words_.size() * 4);
    if (e == hipSuccess) e = hipMemcpy(*block, words_.data(), words_.size() * 4, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      set_error(std::string(what) + " upload: " + hipGetErrorString(e));
      if (*block) (void)hipFree(*block);
      *block = nullptr;
      return WMPC_ERR_HIP;
    }
    base_ = static_cast<const uint32_t*>(*block);
    return WMPC_OK;
  }
  template <class T>
  const T* at(size_t i) const {
    return reinterpret_cast<const T*>(base_ + offs_[i]);
  }

 private:
  std::vector<uint32_t> words_;
  std::vector<size_t> offs_;
  const uint32_t* base_ = nullptr;
};

inline int32_t float_bits(double v) {
  const float f = (float)v;
  int32_t i;
  std::memcpy(&i, &f, 4);
  return i;
}

}  // namespace wmpc
