// CPU stand-ins for the rocPRIM calls of the CSR build and the list sort.
#pragma once
#include <algorithm>
#include <vector>
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
// Stable, like rocPRIM's LSD radix sort: values follow their keys.
template <class K, class V>
hipError_t radix_sort_pairs(void *tmp, size_t &bytes, double_buffer<K> &keys, double_buffer<V> &vals, size_t n, unsigned,
                            unsigned, hipStream_t = nullptr, bool = false) {
  if (!tmp) {
    bytes = 256;
    return hipSuccess;
  }
  std::vector<size_t> idx(n);
  for (size_t i = 0; i < n; ++i) idx[i] = i;
  const K *k = keys.current();
  std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return k[a] < k[b]; });
  K *ko = keys.alternate();
  V *vo = vals.alternate();
  const V *v = vals.current();
  for (size_t i = 0; i < n; ++i) {
    ko[i] = k[idx[i]];
    vo[i] = v[idx[i]];
  }
  keys.sw = !keys.sw;
  vals.sw = !vals.sw;
  return hipSuccess;
}
}  // namespace rocprim
