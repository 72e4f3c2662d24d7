// Fast decimal text emit/scan for the .ckt/.flst/.ptn formats.
#pragma once
#include <cstdint>
#include <cstring>
#include <string>

namespace lsim {

// Appends unsigned decimals to a growing std::string.
class TextOut {
 public:
  explicit TextOut(std::string &dst) : s_(dst) {}
  TextOut &u(uint64_t v) {
    char buf[24];
    int n = 0;
    do {
      buf[n++] = char('0' + v % 10);
      v /= 10;
    } while (v);
    size_t at = s_.size();
    s_.resize(at + n);
    for (int i = 0; i < n; ++i) s_[at + i] = buf[n - 1 - i];
    return *this;
  }
  TextOut &c(char ch) {
    s_.push_back(ch);
    return *this;
  }

 private:
  std::string &s_;
};

// Reads whitespace-separated unsigned decimals the way `std::istream >> size_t`
// does for well-formed input; anything else is reported as an error instead of
// leaving the stream in a failed state.
class TextIn {
 public:
  TextIn(const char *p, size_t n) : p_(p), end_(p + n) {}
  // Returns false at end of input or on a token that is not a plain unsigned
  // decimal fitting in 64 bits; `bad()` distinguishes the two.
  bool next(uint64_t &out) {
    while (p_ < end_ && is_space(*p_)) ++p_;
    if (p_ >= end_) return false;
    if (*p_ < '0' || *p_ > '9') {
      bad_ = true;
      return false;
    }
    uint64_t v = 0;
    while (p_ < end_ && *p_ >= '0' && *p_ <= '9') {
      uint64_t d = uint64_t(*p_ - '0');
      if (v > (UINT64_MAX - d) / 10) {
        bad_ = true;
        return false;
      }
      v = v * 10 + d;
      ++p_;
    }
    if (p_ < end_ && !is_space(*p_)) {
      bad_ = true;
      return false;
    }
    out = v;
    return true;
  }
  bool bad() const { return bad_; }

 private:
  static bool is_space(char c) { return c == ' ' || c == '\n' || c == '\t' || c == '\r' || c == '\f' || c == '\v'; }
  const char *p_;
  const char *end_;
  bool bad_ = false;
};

}  // namespace lsim
