// fg_bal.cpp -- BAL problem text: reader with the reference's validation
// (examples/bal/model.rs:59-113) and a writer whose output the reader returns bit for bit.
//
// Token grammar follows the reference's use of Rust's `str::parse`: whitespace-separated
// tokens; counts and indices are unsigned decimal integers (an optional '+', no '-', no
// overflow); values are decimal floats `[+-]? (d+ [. d*] | . d+) ([eE] [+-]? d+)?` or
// inf / infinity / nan in any case (the latter rejected as nonfinite). Conversion is
// correctly rounded (strtod on the validated token).
#include <cctype>
#include <charconv>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <locale.h>
#include <stdlib.h>
#include <string>

#include "fagra_gpu.h"

namespace {

struct Tokens {
  const char* p;
  const char* end;
  // Rust's split_whitespace on ASCII text: space, \t, \n, \v, \f, \r
  static bool space(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r'; }
  bool next(const char** b, size_t* n) {
    while (p < end && space(*p)) ++p;
    if (p >= end) return false;
    const char* s = p;
    while (p < end && !space(*p)) ++p;
    *b = s;
    *n = (size_t)(p - s);
    return true;
  }
};

bool parse_usize(const char* s, size_t n, size_t* out) {
  size_t i = 0;
  if (i < n && s[i] == '+') ++i;
  if (i == n) return false;
  uint64_t v = 0;
  for (; i < n; ++i) {
    if (s[i] < '0' || s[i] > '9') return false;
    const uint64_t d = (uint64_t)(s[i] - '0');
    if (v > (std::numeric_limits<uint64_t>::max() - d) / 10) return false;
    v = 10 * v + d;
  }
  *out = (size_t)v;
  return true;
}

bool ieq(const char* s, size_t n, const char* word) {
  const size_t m = std::strlen(word);
  if (n != m) return false;
  for (size_t i = 0; i < n; ++i)
    if (std::tolower((unsigned char)s[i]) != word[i]) return false;
  return true;
}

bool parse_f64(const char* s, size_t n, double* out) {
  size_t i = 0;
  bool neg = false;
  if (i < n && (s[i] == '+' || s[i] == '-')) neg = s[i++] == '-';
  const char* body = s + i;
  const size_t m = n - i;
  if (ieq(body, m, "inf") || ieq(body, m, "infinity")) {
    *out = neg ? -std::numeric_limits<double>::infinity() : std::numeric_limits<double>::infinity();
    return true;
  }
  if (ieq(body, m, "nan")) {
    *out = std::numeric_limits<double>::quiet_NaN();
    return true;
  }
  size_t j = 0, digits = 0;
  while (j < m && std::isdigit((unsigned char)body[j])) ++j, ++digits;
  if (j < m && body[j] == '.') {
    ++j;
    while (j < m && std::isdigit((unsigned char)body[j])) ++j, ++digits;
  }
  if (digits == 0) return false;
  if (j < m && (body[j] == 'e' || body[j] == 'E')) {
    ++j;
    if (j < m && (body[j] == '+' || body[j] == '-')) ++j;
    size_t e = 0;
    while (j < m && std::isdigit((unsigned char)body[j])) ++j, ++e;
    if (e == 0) return false;
  }
  if (j != m) return false;
  char small[64];
  std::string big;
  const char* z;
  if (n < sizeof(small)) {
    std::memcpy(small, s, n);
    small[n] = 0;
    z = small;
  } else {
    big.assign(s, n);
    z = big.c_str();
  }
  // "C" numeric locale whatever the process's; ERANGE results (+-inf, subnormals, zero)
  // are the correctly rounded values
  static const locale_t c_locale = newlocale(LC_ALL_MASK, "C", (locale_t)0);
  *out = strtod_l(z, nullptr, c_locale);
  return true;
}

void set_err(char* err, size_t cap, const char* msg) {
  if (err && cap) std::snprintf(err, cap, "%s", msg);
}

// The reference parser. store == false validates the whole text and reports its sizes
// without writing values (callers size their arrays from a text known to be valid);
// store == true requires the reported sizes and fills the arrays.
fg_status parse(const char* text, size_t len, fg_bal_sizes* sizes, bool store, double* cameras, double* points,
                int64_t* obs_camera, int64_t* obs_point, double* pixels, char* err, size_t err_len) {
  Tokens t{text, text + len};
  const char* b;
  size_t n;
  auto fail = [&](const char* msg) {
    set_err(err, err_len, msg);
    return FG_INVALID_INPUT;
  };
  auto next_usize = [&](size_t* v) -> int {
    if (!t.next(&b, &n)) return 1;
    return parse_usize(b, n, v) ? 0 : 2;
  };
  auto next_f64 = [&](double* v) -> int {
    if (!t.next(&b, &n)) return 1;
    return parse_f64(b, n, v) ? 0 : 2;
  };
  static const char* kTrunc = "truncated BAL input";
  static const char* kToken = "invalid BAL token";
  size_t h[3];
  for (size_t& v : h) {
    const int r = next_usize(&v);
    if (r) return fail(r == 1 ? kTrunc : kToken);
  }
  if (h[0] == 0 || h[1] == 0 || h[2] == 0) return fail("empty BAL problem");
  if (store && (h[0] != sizes->n_cameras || h[1] != sizes->n_points || h[2] != sizes->n_observations))
    return FG_DIMENSION_MISMATCH;
  const size_t nc = h[0], np = h[1], no = h[2];
  // 9 nc + 3 np values follow the observations; counts beyond what the text can hold end
  // as truncated input once its tokens run out
  const size_t n_values = nc > SIZE_MAX / 16 || np > SIZE_MAX / 16 ? SIZE_MAX : 9 * nc + 3 * np;
  for (size_t k = 0; k < no; ++k) {
    size_t c, p;
    double u, v;
    int r = next_usize(&c);
    if (!r) r = next_usize(&p);
    if (!r) r = next_f64(&u);
    if (!r) r = next_f64(&v);
    if (r) return fail(r == 1 ? kTrunc : kToken);
    if (c >= nc || p >= np || !std::isfinite(u) || !std::isfinite(v))
      return fail("invalid observation index or value");
    if (store) {
      obs_camera[k] = (int64_t)c;
      obs_point[k] = (int64_t)p;
      pixels[2 * k] = u;
      pixels[2 * k + 1] = v;
    }
  }
  for (size_t i = 0; i < n_values; ++i) {
    double x;
    const int r = next_f64(&x);
    if (r) return fail(r == 1 ? kTrunc : kToken);
    if (!std::isfinite(x)) return fail("nonfinite parameter");
    if (!store) continue;
    if (i < 9 * nc) cameras[i] = x;
    else points[i - 9 * nc] = x;
  }
  if (t.next(&b, &n)) return fail("trailing BAL tokens");
  if (!store) {
    sizes->n_cameras = nc;
    sizes->n_points = np;
    sizes->n_observations = no;
  }
  return FG_OK;
}

// Shortest decimal that reads back to the same double.
size_t put_f64(char* dst, double x) {
  char tmp[32];
  const auto r = std::to_chars(tmp, tmp + sizeof(tmp), x);
  const size_t n = (size_t)(r.ptr - tmp);
  if (dst) std::memcpy(dst, tmp, n);
  return n;
}

size_t put_u64(char* dst, uint64_t x) {
  char tmp[24];
  const auto r = std::to_chars(tmp, tmp + sizeof(tmp), x);
  const size_t n = (size_t)(r.ptr - tmp);
  if (dst) std::memcpy(dst, tmp, n);
  return n;
}

}  // namespace

extern "C" {

fg_status fg_bal_scan(const char* text, size_t len, fg_bal_sizes* sizes, char* err, size_t err_len) {
  if (!text || !sizes) return FG_INVALID_OPTIONS;
  return parse(text, len, sizes, false, nullptr, nullptr, nullptr, nullptr, nullptr, err, err_len);
}

fg_status fg_bal_parse(const char* text, size_t len, const fg_bal_sizes* sizes, double* cameras, double* points,
                       int64_t* obs_camera, int64_t* obs_point, double* pixels, char* err, size_t err_len) {
  if (!text || !sizes || !cameras || !points || !obs_camera || !obs_point || !pixels) return FG_INVALID_OPTIONS;
  fg_bal_sizes s = *sizes;
  return parse(text, len, &s, true, cameras, points, obs_camera, obs_point, pixels, err, err_len);
}

fg_status fg_bal_format(const fg_bal_sizes* sizes, const double* cameras, const double* points,
                        const int64_t* obs_camera, const int64_t* obs_point, const double* pixels, char* buf,
                        size_t cap, size_t* written) {
  if (!sizes || !cameras || !points || !obs_camera || !obs_point || !pixels || !written) return FG_INVALID_OPTIONS;
  const size_t nc = sizes->n_cameras, np = sizes->n_points, no = sizes->n_observations;
  if (nc == 0 || np == 0 || no == 0) return FG_INVALID_OPTIONS;
  for (size_t k = 0; k < no; ++k)
    if (obs_camera[k] < 0 || (size_t)obs_camera[k] >= nc || obs_point[k] < 0 || (size_t)obs_point[k] >= np ||
        !std::isfinite(pixels[2 * k]) || !std::isfinite(pixels[2 * k + 1]))
      return FG_INVALID_OPTIONS;
  for (size_t i = 0; i < 9 * nc; ++i)
    if (!std::isfinite(cameras[i])) return FG_INVALID_OPTIONS;
  for (size_t i = 0; i < 3 * np; ++i)
    if (!std::isfinite(points[i])) return FG_INVALID_OPTIONS;
  // one pass measures (buf == null or too small), the same pass writes when it fits
  size_t at = 0;
  const bool emit = buf != nullptr;
  auto room = [&](size_t n) { return emit && at + n <= cap; };
  char line[96];
  auto flush = [&](size_t n) {
    if (room(n)) std::memcpy(buf + at, line, n);
    at += n;
  };
  size_t n = put_u64(line, nc);
  line[n++] = ' ';
  n += put_u64(line + n, np);
  line[n++] = ' ';
  n += put_u64(line + n, no);
  line[n++] = '\n';
  flush(n);
  for (size_t k = 0; k < no; ++k) {
    n = put_u64(line, (uint64_t)obs_camera[k]);
    line[n++] = ' ';
    n += put_u64(line + n, (uint64_t)obs_point[k]);
    line[n++] = ' ';
    n += put_f64(line + n, pixels[2 * k]);
    line[n++] = ' ';
    n += put_f64(line + n, pixels[2 * k + 1]);
    line[n++] = '\n';
    flush(n);
  }
  for (size_t i = 0; i < 9 * nc + 3 * np; ++i) {
    n = put_f64(line, i < 9 * nc ? cameras[i] : points[i - 9 * nc]);
    line[n++] = '\n';
    flush(n);
  }
  *written = at;
  return emit && at <= cap ? FG_OK : FG_DIMENSION_MISMATCH;
}

}  // extern "C"
