// SUM over f64 cells, exact and order-independent, rounded once.
//
// Reference paths replaced: the AggregateSumF64 kernels (kernel/aggregate.rs:19-147:
// `cells.iter().map(f64::from_bits).sum()` and 4- / 8-lane VADDPD variants) and the
// masked f64 sums of the dispatch path (exec/vectorized.rs:128-136, :525-533,
// engine/dispatch.rs:400-404 for FLOAT / DECIMAL columns).  Those add in an order
// (sequential, or lane-strided then a horizontal sum), so their roundings differ
// from one another; on the device the order of additions is not fixed at all.
// Instead every finite double is added EXACTLY: the value m * 2^(e-1075) is the
// integer m << (e - 1) (subnormals: m) in units of 2^-1074, a number below
// 2^2098, which splits into at most three 32-bit pieces of a fixed-point integer
// kept as 72 limbs of 32 bits.  Each workgroup adds pieces to int64 LDS limbs
// (exact for < 2^31 values per workgroup), normalises its limbs to [0, 2^32)
// plus a signed top limb, and adds them to the device limbs; the host rounds the
// exact total to the nearest double, ties to even.  The result is the correctly
// rounded sum (what math.fsum returns), which equals the reference's sum
// whenever the reference's partial sums are exact (integer-valued cells below
// 2^53, as in its tests).  NaN, +Inf and -Inf follow IEEE addition: any NaN, or
// both infinities, give NaN; otherwise an infinity gives that infinity.
#include <cmath>

#include "tgp_internal.h"

namespace tgp {

namespace {

constexpr int kLimbs = 72;
// flags: 1 NaN, 2 +Inf, 4 -Inf, 8 an addend other than -0.0 (for the sign of zero)
constexpr unsigned kNaN = 1, kPosInf = 2, kNegInf = 4, kNotNegZero = 8;

__global__ void __launch_bounds__(256) k_sum_f64(const uint64_t* __restrict__ cells, uint64_t n,
                                                 const uint64_t* __restrict__ sel,
                                                 long long* __restrict__ limbs,
                                                 unsigned int* __restrict__ flags) {
  __shared__ long long acc[kLimbs];
  __shared__ unsigned int wflags;
  for (int i = threadIdx.x; i < kLimbs; i += blockDim.x) acc[i] = 0;
  if (threadIdx.x == 0) wflags = 0;
  __syncthreads();
  unsigned fl = 0;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    if (sel && !((sel[i >> 6] >> (i & 63)) & 1)) continue;
    const uint64_t b = cells[i];
    const uint32_t e = (uint32_t)(b >> 52) & 0x7ff;
    const uint64_t f = b & ((1ull << 52) - 1);
    const bool neg = (b >> 63) != 0;
    if (e == 0x7ff) {
      fl |= f ? kNaN : (neg ? kNegInf : kPosInf);
      continue;
    }
    if (!(neg && e == 0 && f == 0)) fl |= kNotNegZero;
    if (e == 0 && f == 0) continue;
    const uint64_t m = e ? (f | (1ull << 52)) : f;
    const uint32_t sh = e ? e - 1 : 0;
    const uint32_t j = sh >> 5;
    const unsigned __int128 w = (unsigned __int128)m << (sh & 31);
    const long long p[3] = {(long long)(uint32_t)w, (long long)(uint32_t)(w >> 32),
                            (long long)(uint32_t)(w >> 64)};
#pragma unroll
    for (int k = 0; k < 3; ++k)
      if (p[k]) atomicAdd(reinterpret_cast<unsigned long long*>(&acc[j + k]),
                          (unsigned long long)(neg ? -p[k] : p[k]));
  }
  if (fl) atomicOr(&wflags, fl);
  __syncthreads();
  if (threadIdx.x == 0) {
    // limbs 0 .. kLimbs-2 to [0, 2^32), the carry into the signed top limb
    long long carry = 0;
    for (int j = 0; j < kLimbs - 1; ++j) {
      const long long v = acc[j] + carry;
      acc[j] = v & 0xffffffffll;
      carry = v >> 32;  // arithmetic shift: floor division by 2^32
    }
    acc[kLimbs - 1] += carry;
    if (wflags) atomicOr(flags, wflags);
  }
  __syncthreads();
  for (int j = threadIdx.x; j < kLimbs; j += blockDim.x)
    if (acc[j]) atomicAdd(reinterpret_cast<unsigned long long*>(&limbs[j]),
                          (unsigned long long)acc[j]);
}

// The exact total sum_j limbs[j] * 2^(32 j) * 2^-1074 rounded to the nearest double
// (ties to even).
double round_limbs(const long long* limbs, unsigned fl, bool neg_zero_start) {
  if ((fl & kNaN) || ((fl & kPosInf) && (fl & kNegInf))) return std::nan("");
  if (fl & kPosInf) return HUGE_VAL;
  if (fl & kNegInf) return -HUGE_VAL;
  // two's complement big integer in 32-bit limbs
  uint32_t mag[kLimbs];
  __int128 carry = 0;
  for (int j = 0; j < kLimbs; ++j) {
    const __int128 v = (__int128)limbs[j] + carry;
    mag[j] = (uint32_t)(v & 0xffffffff);
    carry = v >> 32;
  }
  const bool neg = carry < 0;
  if (neg) {  // magnitude = two's complement negation
    uint64_t c = 1;
    for (int j = 0; j < kLimbs; ++j) {
      const uint64_t v = (uint64_t)(uint32_t)~mag[j] + c;
      mag[j] = (uint32_t)v;
      c = v >> 32;
    }
  }
  int top = -1;
  for (int j = kLimbs - 1; j >= 0 && top < 0; --j)
    if (mag[j]) top = 32 * j + 31 - __builtin_clz(mag[j]);
  if (top < 0) return (neg_zero_start && !(fl & kNotNegZero)) ? -0.0 : 0.0;
  auto bit = [&](int k) -> uint64_t { return k < 0 ? 0 : (mag[k >> 5] >> (k & 31)) & 1; };
  double r;
  if (top < 53) {  // below 2^53 units: exactly representable (subnormal or normal)
    uint64_t v = 0;
    for (int k = top; k >= 0; --k) v = (v << 1) | bit(k);
    r = std::ldexp((double)v, -1074);
  } else {
    const int s = top - 52;  // the 53 bits [s, top] form the significand
    uint64_t mant = 0;
    for (int k = top; k >= s; --k) mant = (mant << 1) | bit(k);
    const uint64_t half = bit(s - 1);
    bool sticky = false;
    for (int k = s - 2; k >= 0 && !sticky; --k) sticky = bit(k) != 0;
    if (half && (sticky || (mant & 1))) ++mant;  // 2^53 is still exact as a double
    r = std::ldexp((double)mant, s - 1074);      // >= 2^-1021: normal; overflow -> inf
  }
  return neg ? -r : r;
}

}  // namespace

int32_t sum_f64_exact(tgp_ctx* ctx, const uint64_t* d_cells, uint64_t n,
                      const uint64_t* d_sel_bitmap, bool neg_zero_start, hipStream_t s,
                      double* out) {
  void* scr = nullptr;
  TGP_TRY(get_scratch(ctx, s, kLimbs * 8 + 8, &scr));
  long long* limbs = static_cast<long long*>(scr);
  unsigned int* flags = reinterpret_cast<unsigned int*>(limbs + kLimbs);
  {
    uint64_t* zp[1] = {reinterpret_cast<uint64_t*>(limbs)};
    const uint64_t zn[1] = {kLimbs + 1};
    TGP_TRY(zero_ranges(s, 1, zp, zn));
  }
  if (n) {
    uint64_t g = (n + 255) / 256;
    const uint64_t gmax = (uint64_t)persistent_grid(ctx, 4);
    if (g > gmax) g = gmax;
    k_sum_f64<<<(unsigned)g, 256, 0, s>>>(d_cells, n, d_sel_bitmap, limbs, flags);
    TGP_HIP(hipGetLastError());
  }
  long long h[kLimbs + 1];
  TGP_HIP(hipMemcpyAsync(h, limbs, sizeof(h), hipMemcpyDeviceToHost, s));
  TGP_HIP(hipStreamSynchronize(s));
  unsigned fl;
  std::memcpy(&fl, &h[kLimbs], sizeof(fl));
  *out = round_limbs(h, fl, neg_zero_start);
  return TGP_OK;
}

}  // namespace tgp

using namespace tgp;

extern "C" {

int32_t tgp_sum_f64_sync(tgp_ctx* ctx, const uint64_t* d_cells, uint64_t n,
                         const uint64_t* d_sel_bitmap, int32_t neg_zero_start, double* h_sum,
                         void* stream) {
  TGP_REQUIRE(ctx && h_sum && (n == 0 || d_cells), "tgp_sum_f64_sync: null argument");
  TGP_TRY(check_injected(ctx));
  TGP_REQUIRE((reinterpret_cast<uintptr_t>(d_cells) & 7) == 0,
              "tgp_sum_f64_sync: cells must be 8-byte aligned");
  return sum_f64_exact(ctx, d_cells, n, d_sel_bitmap, neg_zero_start != 0,
                       pick_stream(ctx, stream), h_sum);
}

}  // extern "C"
