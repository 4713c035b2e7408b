// A minimal test harness in the style of googletest (the reference's test
// framework, test/rtspatial_tests.cu:1-13): TEST(suite, name) registers a
// case, ASSERT_* aborts the case with a message, main() runs them all and
// returns non-zero if any failed.
#ifndef RTSPATIAL_AMD_MINI_TEST_H
#define RTSPATIAL_AMD_MINI_TEST_H

#include <cstdio>
#include <functional>
#include <sstream>
#include <string>
#include <vector>

namespace mini_test {

struct Failure {
  std::string msg;
};

struct Case {
  std::string name;
  std::function<void()> fn;
};

inline std::vector<Case>& registry() {
  static std::vector<Case> r;
  return r;
}

struct Registrar {
  Registrar(const char* suite, const char* name, std::function<void()> fn) {
    registry().push_back({std::string(suite) + "." + name, std::move(fn)});
  }
};

inline int run_all(const char* filter) {
  int failed = 0, ran = 0;
  for (auto& c : registry()) {
    if (filter && c.name.find(filter) == std::string::npos) continue;
    ran++;
    std::printf("[ RUN      ] %s\n", c.name.c_str());
    std::fflush(stdout);
    try {
      c.fn();
      std::printf("[       OK ] %s\n", c.name.c_str());
    } catch (const Failure& f) {
      failed++;
      std::printf("[  FAILED  ] %s: %s\n", c.name.c_str(), f.msg.c_str());
    } catch (const std::exception& e) {
      failed++;
      std::printf("[  FAILED  ] %s: exception %s\n", c.name.c_str(), e.what());
    }
    std::fflush(stdout);
  }
  std::printf("[==========] %d tests ran, %d failed\n", ran, failed);
  return failed ? 1 : 0;
}

}  // namespace mini_test

#define TEST(suite, name)                                                            \
  static void suite##_##name##_body();                                               \
  static mini_test::Registrar suite##_##name##_reg(#suite, #name, suite##_##name##_body); \
  static void suite##_##name##_body()

#define ASSERT_EQ(a, b)                                                              \
  do {                                                                               \
    auto va_ = (a);                                                                  \
    auto vb_ = (b);                                                                  \
    if (!(va_ == vb_)) {                                                             \
      std::ostringstream os_;                                                        \
      os_ << __FILE__ << ":" << __LINE__ << ": " #a " == " #b " (" << va_ << " vs " << vb_ << ")"; \
      throw mini_test::Failure{os_.str()};                                           \
    }                                                                                \
  } while (0)

#define ASSERT_TRUE(c)                                                               \
  do {                                                                               \
    if (!(c)) {                                                                      \
      std::ostringstream os_;                                                        \
      os_ << __FILE__ << ":" << __LINE__ << ": " #c;                                 \
      throw mini_test::Failure{os_.str()};                                           \
    }                                                                                \
  } while (0)

#endif
