// Result encoding for exact int128 values (host code).
//
// The reference's result cells are u64 (engine/result.rs:37-53; Value2::to_u64,
// query_interpreter/types.rs:383-391), and pgwire prints a cell as the decimal
// text of the u64 (server/async_pgwire.rs:881-910), so an int128 sum has no
// carrier.  ResultColumn already carries per-row strings (`string_values`) and
// a type OID: the exact sum travels as NUMERIC text (OID 1700), scaled by the
// column's decimal scale (cents: 2).  An exact AVG is the int128 sum over the
// count, rounded half away from zero to a chosen number of fraction digits.
#include "tgp_internal.h"

#include <cstring>
#include <string>

namespace tgp {

typedef unsigned __int128 u128;
typedef __int128 i128;

// Decimal digits of a magnitude, most significant first.
static std::string digits_of(u128 m) {
  if (m == 0) return "0";
  char tmp[48];
  int k = 0;
  while (m) {
    tmp[k++] = (char)('0' + (int)(m % 10));
    m /= 10;
  }
  std::string s(k, '0');
  for (int i = 0; i < k; ++i) s[i] = tmp[k - 1 - i];
  return s;
}

// "[-]int[.frac]" of negative ? -mag : mag, scaled by 10^-scale.
static std::string numeric_text(bool negative, const std::string& mag, int32_t scale) {
  std::string d = mag;
  if (scale > 0 && (int32_t)d.size() <= scale) d.insert(0, (size_t)(scale + 1 - d.size()), '0');
  std::string out;
  if (negative && d.find_first_not_of('0') != std::string::npos) out.push_back('-');
  if (scale > 0) {
    out.append(d, 0, d.size() - (size_t)scale);
    out.push_back('.');
    out.append(d, d.size() - (size_t)scale, std::string::npos);
  } else {
    out.append(d);
  }
  return out;
}

static int32_t put(const std::string& s, char* buf, uint64_t buf_size, uint64_t* out_len) {
  if (out_len) *out_len = s.size();
  TGP_REQUIRE(buf && buf_size > s.size(), "numeric text needs %llu bytes with its terminator",
              (unsigned long long)s.size() + 1);
  std::memcpy(buf, s.c_str(), s.size() + 1);
  return TGP_OK;
}

}  // namespace tgp

using namespace tgp;

extern "C" {

int32_t tgp_int128_to_numeric(uint64_t lo, uint64_t hi, int32_t scale, char* buf,
                              uint64_t buf_size, uint64_t* out_len) {
  TGP_REQUIRE(scale >= 0 && scale <= 38, "scale %d not in 0..38", scale);
  const i128 v = (i128)(((u128)hi << 64) | lo);
  const bool neg = v < 0;
  const u128 m = neg ? (u128)0 - (u128)v : (u128)v;
  return put(numeric_text(neg, digits_of(m), scale), buf, buf_size, out_len);
}

int32_t tgp_int128_avg_to_numeric(uint64_t lo, uint64_t hi, uint64_t count, int32_t scale,
                                  int32_t frac_digits, char* buf, uint64_t buf_size,
                                  uint64_t* out_len) {
  TGP_REQUIRE(count > 0, "AVG of no rows");
  TGP_REQUIRE(scale >= 0 && scale <= 38 && frac_digits >= 0 && frac_digits <= 38 &&
                  frac_digits >= scale,
              "scale %d, fraction digits %d: need 0 <= scale <= digits <= 38", scale,
              frac_digits);
  const i128 v = (i128)(((u128)hi << 64) | lo);
  const bool neg = v < 0;
  const u128 m = neg ? (u128)0 - (u128)v : (u128)v;
  // |v| / count as a decimal with frac_digits fraction digits (the value's own
  // scale already holds `scale` of them): long division digit by digit, so no
  // intermediate exceeds 128 bits (remainders stay below count < 2^64)
  u128 q = m / count, r = m % count;
  std::string mag = digits_of(q);
  for (int32_t i = scale; i < frac_digits; ++i) {
    r *= 10;
    mag.push_back((char)('0' + (int)(r / count)));
    r %= count;
  }
  // round half away from zero on the remainder
  if (2 * r >= (u128)count) {
    int k = (int)mag.size() - 1;
    while (k >= 0 && mag[k] == '9') mag[k--] = '0';
    if (k < 0) mag.insert(mag.begin(), '1');
    else mag[k] = (char)(mag[k] + 1);
  }
  size_t nz = mag.find_first_not_of('0');
  mag = nz == std::string::npos ? "0" : mag.substr(nz);
  return put(numeric_text(neg, mag, frac_digits), buf, buf_size, out_len);
}

}  // extern "C"
