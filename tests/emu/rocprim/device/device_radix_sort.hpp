// CPU stand-in for the one rocPRIM call the CSR build makes.
#pragma once
#include <algorithm>
#include "hip/hip_runtime.h"
namespace rocprim {
template <class T>
struct double_buffer {
  T *a, *b;
  bool sw = false;
  double_buffer(T *x, T *y) : a(x), b(y) {}
  T *current() const { return sw ? b : a; }
  T *alternate() const { return sw ? a : b; }
};
template <class K>
hipError_t radix_sort_keys(void *tmp, size_t &bytes, double_buffer<K> &db, size_t n, unsigned, unsigned, hipStream_t = nullptr,
                           bool = false) {
  if (!tmp) {
    bytes = 256;
    return hipSuccess;
  }
  std::sort(db.current(), db.current() + n);
  return hipSuccess;
}
}  // namespace rocprim
