// xb_sort.hip -- device radix sort of 64-bit keys (rocPRIM), kept in a
// translation unit of its own so the sort's instantiation compiles once.
#include <rocprim/device/device_radix_sort.hpp>

#include <stdexcept>
#include <string>

#include "xb_internal.hpp"

namespace xb {

size_t sort_u64_temp_bytes(int32_t n) {
  size_t bytes = 0;
  const hipError_t e = rocprim::radix_sort_keys(nullptr, bytes, (const unsigned long long*)nullptr,
                                                (unsigned long long*)nullptr, (size_t)(n > 0 ? n : 1), 0, 64);
  if (e != hipSuccess) throw std::runtime_error(std::string("radix sort size query: ") + hipGetErrorString(e));
  return bytes;
}

void sort_u64(void* temp, size_t bytes, const unsigned long long* in, unsigned long long* out, int32_t n,
              hipStream_t st) {
  if (n <= 0) return;
  const hipError_t e = rocprim::radix_sort_keys(temp, bytes, in, out, (size_t)n, 0, 64, st);
  if (e != hipSuccess) throw std::runtime_error(std::string("radix sort: ") + hipGetErrorString(e));
}

}  // namespace xb
