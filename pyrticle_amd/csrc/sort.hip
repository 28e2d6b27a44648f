// sort.hip -- stable radix sort of (key, index) pairs for the element sorts of
// the deposition, the adjoint pushes and the work-item order of the list
// gathers, on rocPRIM's device radix sort (its own translation unit: the
// library's headers are large).  Keys are clamped to kmax on the fly, so that
// element ids of particles outside the mesh (all ones) sort last, and only
// the bits of kmax are sorted.  The sort is stable: equal keys keep their
// index order, so the permutation is the one numpy.argsort(kind="stable")
// gives and every result built on it is reproducible.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/iterator/counting_iterator.hpp>
#include <rocprim/iterator/transform_iterator.hpp>

#include "pyr_internal.h"

namespace pyr {

namespace {
struct ClampKey {
  uint32_t kmax;
  __host__ __device__ uint32_t operator()(uint32_t k) const { return k < kmax ? k : kmax; }
};

int key_bits(uint32_t kmax) {
  int bits = 1;
  while (bits < 32 && ((uint64_t)1 << bits) <= (uint64_t)kmax) ++bits;
  return bits;
}
}  // namespace

const uint32_t *sort_pairs(DevBuf &scratch, int64_t n, const uint32_t *keys, uint32_t kmax,
                           uint32_t *kout, uint32_t *vout, hipStream_t s) {
  if (n <= 0) return vout;
  if (n > (int64_t)UINT32_MAX) throw std::runtime_error("sort_pairs: more than 2^32 - 1 keys");
  const unsigned bits = (unsigned)key_bits(kmax);
  auto kin = rocprim::make_transform_iterator(keys, ClampKey{kmax});
  auto vin = rocprim::make_counting_iterator<uint32_t>(0u);
  const uint32_t size = (uint32_t)n;
  size_t bytes = 0;
  PYR_HIP(rocprim::radix_sort_pairs(nullptr, bytes, kin, kout, vin, vout, size, 0u, bits, s));
  scratch.reserve(bytes);
  PYR_HIP(rocprim::radix_sort_pairs(scratch.ptr, bytes, kin, kout, vin, vout, size, 0u, bits, s));
  return vout;
}

}  // namespace pyr
