// textnum.h — number parsing of the model-text reader (plain C++, no device
// code), shared by the engine and by the CPU test that pins it against the
// reference's known answers (fmt.rs:160-273).
#ifndef LGBMHIP_TEXTNUM_H_
#define LGBMHIP_TEXTNUM_H_

#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <limits>
#include <string>
#include <vector>

// fmt.rs:57-64 parse_f64 (Rust's f64 grammar, correctly rounded, plus MSVC's
// "nan(ind)"): [+-] then inf / infinity / nan (any case), or digits with an
// optional fraction and exponent. glibc's strtod rounds correctly, so the
// token is checked against that grammar and converted by strtod.
inline bool parse_f64(const std::string& tok, double* out) {
  if (tok == "nan(ind)" || tok == "-nan(ind)") {
    *out = std::numeric_limits<double>::quiet_NaN();
    return true;
  }
  size_t i = 0;
  bool neg = false;
  if (i < tok.size() && (tok[i] == '+' || tok[i] == '-')) neg = tok[i++] == '-';
  std::string rest;
  for (size_t k = i; k < tok.size(); ++k) rest += (char)std::tolower((unsigned char)tok[k]);
  if (rest == "inf" || rest == "infinity") {
    *out = neg ? -std::numeric_limits<double>::infinity() : std::numeric_limits<double>::infinity();
    return true;
  }
  if (rest == "nan") {
    *out = std::copysign(std::numeric_limits<double>::quiet_NaN(), neg ? -1.0 : 1.0);
    return true;
  }
  size_t k = i, digits = 0;
  while (k < tok.size() && std::isdigit((unsigned char)tok[k])) ++k, ++digits;
  if (k < tok.size() && tok[k] == '.') {
    ++k;
    while (k < tok.size() && std::isdigit((unsigned char)tok[k])) ++k, ++digits;
  }
  if (digits == 0) return false;
  if (k < tok.size() && (tok[k] == 'e' || tok[k] == 'E')) {
    ++k;
    if (k < tok.size() && (tok[k] == '+' || tok[k] == '-')) ++k;
    size_t ed = 0;
    while (k < tok.size() && std::isdigit((unsigned char)tok[k])) ++k, ++ed;
    if (ed == 0) return false;
  }
  if (k != tok.size()) return false;
  *out = std::strtod(tok.c_str(), nullptr);
  return true;
}

// base^power with the multiplication order of upstream's Common::Pow
// (fmt.rs:66-79), which sets the legacy parser's rounding.
inline double pow_legacy(double base, int power) {
  if (power < 0) return 1.0 / pow_legacy(base, -power);
  if (power == 0) return 1.0;
  if (power % 2 == 0) return pow_legacy(base * base, power / 2);
  if (power % 3 == 0) return pow_legacy(base * base * base, power / 3);
  return base * pow_legacy(base, power - 1);
}

// fmt.rs:81-150 atof_legacy: upstream's Common::Atof, the parser of the
// split_gain, internal_value, internal_weight and shrinkage fields. It is not
// correctly rounded; its digit loop and power-of-ten scaling give the doubles
// upstream loads. inf / infinity read as +-1e308.
inline bool atof_legacy(const std::string& tok_in, double* out) {
  size_t a = 0, b = tok_in.size();
  while (a < b && tok_in[a] == ' ') ++a;
  while (b > a && tok_in[b - 1] == ' ') --b;
  const std::string t = tok_in.substr(a, b - a);
  size_t p = 0;
  double sign = 1.0;
  if (p < t.size() && t[p] == '-') sign = -1.0, ++p;
  else if (p < t.size() && t[p] == '+') ++p;
  const bool is_num = p < t.size() && (std::isdigit((unsigned char)t[p]) || t[p] == '.' || t[p] == 'e' || t[p] == 'E');
  if (!is_num) {
    std::string w;
    for (size_t k = p; k < t.size(); ++k) w += (char)std::tolower((unsigned char)t[k]);
    if (w == "na" || w == "nan" || w == "null") { *out = std::numeric_limits<double>::quiet_NaN(); return true; }
    if (w == "inf" || w == "infinity") { *out = sign * 1e308; return true; }
    return false;
  }
  double value = 0.0;
  while (p < t.size() && std::isdigit((unsigned char)t[p])) value = value * 10.0 + (double)(t[p++] - '0');
  if (p < t.size() && t[p] == '.') {
    double right = 0.0;
    int nn = 0;
    ++p;
    while (p < t.size() && std::isdigit((unsigned char)t[p])) right = (double)(t[p++] - '0') + right * 10.0, ++nn;
    value += right / pow_legacy(10.0, nn);
  }
  bool frac = false;
  double scale = 1.0;
  if (p < t.size() && (t[p] == 'e' || t[p] == 'E')) {
    ++p;
    if (p < t.size() && t[p] == '-') frac = true, ++p;
    else if (p < t.size() && t[p] == '+') ++p;
    uint32_t expon = 0;
    while (p < t.size() && std::isdigit((unsigned char)t[p])) expon = expon * 10u + (uint32_t)(t[p++] - '0');
    expon = std::min<uint32_t>(expon, 308);
    for (; expon >= 50; expon -= 50) scale *= 1e50;
    for (; expon >= 8; expon -= 8) scale *= 1e8;
    for (; expon > 0; --expon) scale *= 10.0;
  }
  if (p != t.size()) return false;
  *out = sign * (frac ? value / scale : value * scale);
  return true;
}

// A decimal integer token in [lo, hi] (Rust's str::parse of the field's integer type).
inline bool parse_int(const std::string& tok, long long lo, long long hi, long long* out) {
  size_t i = 0;
  bool neg = false;
  if (i < tok.size() && (tok[i] == '+' || tok[i] == '-')) neg = tok[i++] == '-';
  if (i == tok.size()) return false;
  long long v = 0;
  for (; i < tok.size(); ++i) {
    if (!std::isdigit((unsigned char)tok[i])) return false;
    v = v * 10 + (tok[i] - '0');
    if (v > (long long)1 << 40) return false;
  }
  v = neg ? -v : v;
  if (v < lo || v > hi) return false;
  *out = v;
  return true;
}

inline std::vector<std::string> split_ws(const std::string& s) {
  std::vector<std::string> out;
  std::string cur;
  for (char ch : s) {
    if (ch == ' ' || ch == '\t') {
      if (!cur.empty()) out.push_back(cur);
      cur.clear();
    } else {
      cur += ch;
    }
  }
  if (!cur.empty()) out.push_back(cur);
  return out;
}

#endif  // LGBMHIP_TEXTNUM_H_
