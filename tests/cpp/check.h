// check.h -- minimal test harness for the C++ host-API tests.
#pragma once

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

namespace check {

struct Case {
  const char* name;
  std::function<void()> fn;
};
inline std::vector<Case>& registry() {
  static std::vector<Case> r;
  return r;
}
inline int& failures() {
  static int f = 0;
  return f;
}
struct Reg {
  Reg(const char* n, std::function<void()> f) { registry().push_back({n, f}); }
};

inline bool near_ulps(double a, double b, int ulps) {
  if (a == b) return true;
  const double m = std::fmax(std::fabs(a), std::fabs(b));
  return std::fabs(a - b) <= ulps * (std::nextafter(m, INFINITY) - m);
}

inline int run(int argc, char** argv) {
  const char* only = argc > 1 ? argv[1] : nullptr;
  int ran = 0;
  for (auto& c : registry()) {
    if (only && std::strcmp(only, c.name) != 0) continue;
    const int before = failures();
    c.fn();
    ++ran;
    std::printf("%s %s\n", failures() == before ? "PASS" : "FAIL", c.name);
  }
  std::printf("%d cases, %d failed checks\n", ran, failures());
  return failures() ? 1 : 0;
}

}  // namespace check

#define TEST(name)                                      \
  static void name();                                   \
  static check::Reg reg_##name(#name, name);            \
  static void name()

#define EXPECT_TRUE(c)                                                          \
  do {                                                                          \
    if (!(c)) {                                                                 \
      std::printf("  %s:%d: expected %s\n", __FILE__, __LINE__, #c);           \
      ++check::failures();                                                      \
    }                                                                           \
  } while (0)
#define EXPECT_EQ(a, b) EXPECT_TRUE((a) == (b))
#define EXPECT_NEAR(a, b, tol)                                                  \
  do {                                                                          \
    const double a_ = (a), b_ = (b);                                            \
    if (!(std::fabs(a_ - b_) <= (tol))) {                                       \
      std::printf("  %s:%d: |%.17g - %.17g| > %g\n", __FILE__, __LINE__, a_, b_, \
                  (double)(tol));                                               \
      ++check::failures();                                                      \
    }                                                                           \
  } while (0)
#define EXPECT_DOUBLE_EQ(a, b)                                                  \
  do {                                                                          \
    const double a_ = (a), b_ = (b);                                            \
    if (!check::near_ulps(a_, b_, 4)) {                                         \
      std::printf("  %s:%d: %.17g != %.17g (4 ulps)\n", __FILE__, __LINE__, a_, b_); \
      ++check::failures();                                                      \
    }                                                                           \
  } while (0)
