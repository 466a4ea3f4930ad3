// Bucket hash of the owner engine's LDS tables, shared by device code and
// the host-side range test. Multiplicative (Fibonacci) hashing on a 24-bit
// word: the low 24 bits of key * floor(2^24 / golden ratio), top lb of those
// 24 bits. Consecutive keys land about 0.618 of the table apart, so dense id
// ranges spread evenly even in small tables. On gfx950 this is one full-rate
// 24x24-bit multiply and one bit-field extract (a 32-bit multiply is quarter
// rate). Keys that differ only above bit 23 share a bucket, which costs spills
// only in chunks spanning more than 2^24 ids. Unsigned throughout; the result
// is below 2^lb for 0 <= lb <= 24.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define JC_HD __host__ __device__ __forceinline__
#else
#define JC_HD inline
#endif

namespace jc {

constexpr uint32_t kHashMul24 = 0x9E3779u;  // floor(2^24 * 0.6180339887), odd

// Low 32 bits of the product of the low 24 bits of a and b.
// Only the low 24 bits of the product are used, and from the plain C
// expression the compiler derives that the input masks are redundant and
// emits the quarter-rate 32-bit multiply; the device path names the 24-bit
// multiply explicitly (a register-only ALU instruction).
JC_HD uint32_t mul_u24_hash(uint32_t a) {
#if defined(__HIP_DEVICE_COMPILE__)
  uint32_t r;
  asm("v_mul_u32_u24 %0, 0x9e3779, %1" : "=v"(r) : "v"(a));
  return r;
#else
  return (a & 0xFFFFFFu) * kHashMul24;
#endif
}

// The 24-bit hash word of a key.
JC_HD uint32_t hash24(int32_t y) { return mul_u24_hash((uint32_t)y) & 0xFFFFFFu; }

// Bits [24 - lb, 24) of the hash word; on the device one bit-field extract.
JC_HD uint32_t hash_bucket(int32_t y, int lb) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_ubfe(mul_u24_hash((uint32_t)y), 24 - lb, lb);
#else
  return hash24(y) >> (24 - lb);
#endif
}

}  // namespace jc
