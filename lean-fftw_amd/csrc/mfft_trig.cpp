// Exact-integer twiddle factors, restating the reference's twiddle routine
// LeanFFTW/Core/Trig.lean (whole file) and the pass tables of LeanFFTW/Core/FastPlan.lean.
//
// cos and sin of 2*pi*k/n: exact octant reduction 8k = oct*n + rem (Trig.lean:104-122),
// Taylor series of the reduced angle in 160-bit fixed point (:43-62), one rounding to the
// nearest binary64, ties to even (:67-86). The arithmetic uses unsigned 64-bit limbs;
// every operation is exact integer arithmetic, so the results are the reference's bits.
#include "mfft_internal.h"

#include <cstring>
#include <list>
#include <memory>
#include <mutex>
#include <numeric>
#include <thread>

namespace mfft {
namespace {

constexpr int PREC = 160;
constexpr int LIMBS = 6;  // 384 bits: enough for every product below (< 2^330)

struct Big {
  uint64_t w[LIMBS] = {0, 0, 0, 0, 0, 0};
};

Big from_u64(uint64_t v) {
  Big r;
  r.w[0] = v;
  return r;
}

Big add(const Big &a, const Big &b) {
  Big r;
  unsigned __int128 c = 0;
  for (int i = 0; i < LIMBS; ++i) {
    c += (unsigned __int128)a.w[i] + b.w[i];
    r.w[i] = (uint64_t)c;
    c >>= 64;
  }
  return r;
}

// a - b, assuming a >= b; saturates at zero otherwise (Lean's Nat subtraction)
bool less(const Big &a, const Big &b) {
  for (int i = LIMBS - 1; i >= 0; --i)
    if (a.w[i] != b.w[i]) return a.w[i] < b.w[i];
  return false;
}

Big sub(const Big &a, const Big &b) {
  if (less(a, b)) return Big{};
  Big r;
  uint64_t borrow = 0;
  for (int i = 0; i < LIMBS; ++i) {
    const unsigned __int128 d = (unsigned __int128)a.w[i] - b.w[i] - borrow;
    r.w[i] = (uint64_t)d;
    borrow = (uint64_t)(d >> 64) ? 1 : 0;
  }
  return r;
}

// full product truncated to LIMBS limbs (callers keep products below 2^384)
Big mul(const Big &a, const Big &b) {
  uint64_t t[2 * LIMBS] = {};
  for (int i = 0; i < LIMBS; ++i) {
    if (!a.w[i]) continue;
    unsigned __int128 c = 0;
    for (int j = 0; j + i < 2 * LIMBS && j < LIMBS; ++j) {
      c += (unsigned __int128)a.w[i] * b.w[j] + t[i + j];
      t[i + j] = (uint64_t)c;
      c >>= 64;
    }
  }
  Big r;
  std::memcpy(r.w, t, sizeof r.w);
  return r;
}

Big shr(const Big &a, int s) {
  Big r;
  const int q = s / 64, b = s % 64;
  for (int i = 0; i < LIMBS; ++i) {
    const int k = i + q;
    if (k >= LIMBS) break;
    uint64_t v = a.w[k] >> b;
    if (b && k + 1 < LIMBS) v |= a.w[k + 1] << (64 - b);
    r.w[i] = v;
  }
  return r;
}

Big shl(const Big &a, int s) {
  Big r;
  const int q = s / 64, b = s % 64;
  for (int i = LIMBS - 1; i >= 0; --i) {
    const int k = i - q;
    if (k < 0) continue;
    uint64_t v = a.w[k] << b;
    if (b && k - 1 >= 0) v |= a.w[k - 1] >> (64 - b);
    r.w[i] = v;
  }
  return r;
}

Big div_small(const Big &a, uint64_t d) {
  Big r;
  unsigned __int128 rem = 0;
  for (int i = LIMBS - 1; i >= 0; --i) {
    const unsigned __int128 cur = (rem << 64) | a.w[i];
    r.w[i] = (uint64_t)(cur / d);
    rem = cur % d;
  }
  return r;
}

int bit_length(const Big &a) {
  for (int i = LIMBS - 1; i >= 0; --i)
    if (a.w[i]) return 64 * i + 64 - __builtin_clzll(a.w[i]);
  return 0;
}

bool is_zero(const Big &a) { return bit_length(a) == 0; }

bool bit(const Big &a, int i) { return (a.w[i / 64] >> (i % 64)) & 1; }

// a / b by shift-subtract long division (only for tan's quotient)
Big div_big(const Big &a, const Big &b) {
  Big q, r;
  for (int i = bit_length(a) - 1; i >= 0; --i) {
    r = shl(r, 1);
    if (bit(a, i)) r.w[0] |= 1;
    if (!less(r, b)) {
      r = sub(r, b);
      q.w[i / 64] |= uint64_t(1) << (i % 64);
    }
  }
  return q;
}

// pi * 2^160 from 61 decimal digits, Trig.lean:28-33: piDigits * 2^160 / 10^60
const Big &pi_fixed() {
  static const Big v = [] {
    // piDigits = 3141592653589793238462643383279502884197169399375105820974944
    const char *digits = "3141592653589793238462643383279502884197169399375105820974944";
    Big x;
    for (const char *p = digits; *p; ++p) x = add(mul(x, from_u64(10)), from_u64(uint64_t(*p - '0')));
    x = shl(x, PREC);
    for (int i = 0; i < 60; ++i) x = div_small(x, 10);
    return x;
  }();
  return v;
}

Big fmul(const Big &a, const Big &b) { return shr(mul(a, b), PREC); }

// theta * 2^160 for theta = 2 pi num / den (Trig.lean:39)
Big angle_fixed(uint64_t num, uint64_t den) {
  Big x = mul(shl(pi_fixed(), 1), from_u64(num));
  // den may exceed 2^64 / 2 only for n > 2^60, far outside the domain
  return div_small(x, den);
}

// Values of the series are below 2^162: three 64-bit limbs suffice for every operand.
struct V3 {
  uint64_t w[3];
};

inline V3 v3(const Big &b) { return {{b.w[0], b.w[1], b.w[2]}}; }

inline Big big(const V3 &v) {
  Big b;
  b.w[0] = v.w[0];
  b.w[1] = v.w[1];
  b.w[2] = v.w[2];
  return b;
}

inline bool v3_zero(const V3 &a) { return !(a.w[0] | a.w[1] | a.w[2]); }

// floor(a * b / 2^160)
inline V3 fmul3(const V3 &a, const V3 &b) {
  uint64_t t[6] = {0, 0, 0, 0, 0, 0};
  for (int i = 0; i < 3; ++i) {
    unsigned __int128 c = 0;
    for (int j = 0; j < 3; ++j) {
      c += (unsigned __int128)a.w[i] * b.w[j] + t[i + j];
      t[i + j] = (uint64_t)c;
      c >>= 64;
    }
    t[i + 3] = (uint64_t)c;
  }
  // shift right by 160 = 2 limbs + 32 bits
  V3 r;
  r.w[0] = (t[2] >> 32) | (t[3] << 32);
  r.w[1] = (t[3] >> 32) | (t[4] << 32);
  r.w[2] = (t[4] >> 32) | (t[5] << 32);
  return r;
}

inline V3 div3(const V3 &a, uint64_t d) {
  V3 r;
  unsigned __int128 rem = 0;
  for (int i = 2; i >= 0; --i) {
    const unsigned __int128 cur = (rem << 64) | a.w[i];
    r.w[i] = (uint64_t)(cur / d);
    rem = cur % d;
  }
  return r;
}

inline V3 add3(const V3 &a, const V3 &b) {
  V3 r;
  unsigned __int128 c = 0;
  for (int i = 0; i < 3; ++i) {
    c += (unsigned __int128)a.w[i] + b.w[i];
    r.w[i] = (uint64_t)c;
    c >>= 64;
  }
  return r;
}

// Trig.lean:43-62. Once both terms have been truncated to zero every later term is zero
// too (each is the previous one times a factor, floored), so stopping there adds the same
// zeros the full 39-term loop would.
void cos_sin_fixed(const Big &theta_b, Big &c, Big &s) {
  const V3 theta = v3(theta_b);
  const V3 t2 = fmul3(theta, theta);
  V3 cpos = {{0, 0, uint64_t(1) << 32}}, cneg = {{0, 0, 0}}, spos = theta, sneg = {{0, 0, 0}};
  V3 tc = cpos, ts = theta;
  for (uint64_t j = 1; j < 40; ++j) {
    tc = div3(fmul3(tc, t2), (2 * j - 1) * (2 * j));
    ts = div3(fmul3(ts, t2), (2 * j) * (2 * j + 1));
    if (v3_zero(tc) && v3_zero(ts)) break;
    if (j % 2 == 1) {
      cneg = add3(cneg, tc);
      sneg = add3(sneg, ts);
    } else {
      cpos = add3(cpos, tc);
      spos = add3(spos, ts);
    }
  }
  c = sub(big(cpos), big(cneg));
  s = sub(big(spos), big(sneg));
}

double fixed_to_double(const Big &v) {
  if (is_zero(v)) return 0.0;
  int lg = bit_length(v) - 1;
  uint64_t q;
  if (lg <= 52) {
    q = shl(v, 52 - lg).w[0];
  } else {
    const int sh = lg - 52;
    const Big qq = shr(v, sh);
    q = qq.w[0];
    // remainder vs half: compare the low sh bits
    const Big r = sub(v, shl(qq, sh));
    const Big half = shl(from_u64(1), sh - 1);
    if (less(half, r) || (!less(r, half) && (q & 1))) q += 1;
  }
  if (q == (uint64_t(1) << 53)) {
    q = uint64_t(1) << 52;
    lg += 1;
  }
  const int64_t e = (int64_t)lg - PREC + 1023;
  const uint64_t bits = ((uint64_t)e << 52) | (q - (uint64_t(1) << 52));
  double d;
  std::memcpy(&d, &bits, 8);
  return d;
}

inline double neg_z(double x) { return x == 0.0 ? 0.0 : -x; }

struct CS {
  double c, s;
};

CS small_at(uint64_t n, uint64_t num) {
  Big c, s;
  cos_sin_fixed(angle_fixed(num, 8 * n), c, s);
  return {fixed_to_double(c), fixed_to_double(s)};
}

CS octant(CS cs, uint64_t oct) {
  const double c = cs.c, s = cs.s;
  switch (oct) {
    case 0: return {c, s};
    case 1: return {s, c};
    case 2: return {neg_z(s), c};
    case 3: return {neg_z(c), s};
    case 4: return {neg_z(c), neg_z(s)};
    case 5: return {neg_z(s), neg_z(c)};
    case 6: return {s, neg_z(c)};
    default: return {c, neg_z(s)};
  }
}

}  // namespace

void cos_sin_2pi(uint64_t k, uint64_t n, double &c, double &s) {
  if (n == 0) {
    c = 1.0;
    s = 0.0;
    return;
  }
  k %= n;
  const unsigned __int128 k8 = (unsigned __int128)8 * k;
  const uint64_t oct = (uint64_t)(k8 / n), rem = (uint64_t)(k8 % n);
  const CS r = octant(small_at(n, oct % 2 == 0 ? rem : n - rem), oct);
  c = r.c;
  s = r.s;
}

double tan_2pi(uint64_t num, uint64_t den) {
  Big c, s;
  cos_sin_fixed(angle_fixed(num, den), c, s);
  return fixed_to_double(div_big(shl(s, PREC), c));
}

// Trig.lean:131-134 -- every reduced angle is a multiple of g = gcd(8, n): evaluate the
// n/g + 1 of them once, then map each m < n through the octant symmetries.
namespace {

// The last few tables (sizes up to 2^20), so that planning many candidate schedules of one
// size (the measuring planner) computes its table once.
std::mutex g_table_mu;
std::list<std::pair<uint64_t, std::shared_ptr<const std::vector<double>>>> g_tables;
constexpr size_t kTableCache = 4;

}  // namespace

void cos_sin_2pi_table(uint64_t n, std::vector<double> &cs) {
  const bool cacheable = n <= (uint64_t(1) << 20);
  if (cacheable) {
    std::lock_guard<std::mutex> lk(g_table_mu);
    for (auto it = g_tables.begin(); it != g_tables.end(); ++it)
      if (it->first == n) {
        cs = *it->second;
        g_tables.splice(g_tables.begin(), g_tables, it);  // most recent first
        return;
      }
  }
  cs.assign(2 * n, 0.0);
  const uint64_t g = std::gcd<uint64_t>(8, n);
  const uint64_t count = n / g + 1;
  std::vector<CS> memo(count);
  unsigned nt = std::thread::hardware_concurrency();
  nt = nt == 0 ? 1 : (nt > 16 ? 16 : nt);
  if (count < 4096) nt = 1;
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < nt; ++t)
    pool.emplace_back([&, t] {
      for (uint64_t i = t; i < count; i += nt) memo[i] = small_at(n, g * i);
    });
  for (auto &th : pool) th.join();
  for (uint64_t m = 0; m < n; ++m) {
    const uint64_t oct = 8 * m / n, rem = 8 * m % n;
    const CS r = octant(memo[(oct % 2 == 0 ? rem : n - rem) / g], oct);
    cs[2 * m] = r.c;
    cs[2 * m + 1] = r.s;
  }
  if (cacheable) {
    auto t = std::make_shared<const std::vector<double>>(cs);
    std::lock_guard<std::mutex> lk(g_table_mu);
    g_tables.emplace_front(n, std::move(t));
    if (g_tables.size() > kTableCache) g_tables.pop_back();
  }
}

}  // namespace mfft
