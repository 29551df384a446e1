// Minimal test harness for the C++ facade tests: TEST(group, name) registers
// a case, EXPECT / EXPECT_DOUBLES_EQUAL count failures, main() runs every
// case (or those whose "group.name" contains argv[1]) and exits non-zero on
// any failure.  DATASET_PATH is the folder of the g2o fixtures.
#pragma once

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <exception>
#include <string>
#include <vector>

namespace rpgo_test {

struct Case {
  const char* group;
  const char* name;
  void (*fn)(int&);
};

inline std::vector<Case>& registry() {
  static std::vector<Case> r;
  return r;
}

inline bool add(const char* g, const char* n, void (*fn)(int&)) {
  registry().push_back({g, n, fn});
  return true;
}

inline std::string dataset_path() {
  const char* e = std::getenv("RPGO_DATASET_PATH");
  if (e && *e) return e;
  const char* root = std::getenv("GRAFT_REPO_ROOT");
  return std::string(root && *root ? root : ".") + "/tests/golden/data";
}

}  // namespace rpgo_test

#define DATASET_PATH (rpgo_test::dataset_path())

#define TEST(G, N)                                                                  \
  static void G##_##N##_body(int& rpgo_failures_);                                  \
  static const bool G##_##N##_registered = rpgo_test::add(#G, #N, &G##_##N##_body); \
  static void G##_##N##_body(int& rpgo_failures_)

#define EXPECT(cond)                                                          \
  do {                                                                        \
    if (!(cond)) {                                                            \
      ++rpgo_failures_;                                                       \
      std::fprintf(stderr, "%s:%d: EXPECT(%s) failed\n", __FILE__, __LINE__, #cond); \
    }                                                                         \
  } while (0)

#define EXPECT_DOUBLES_EQUAL(expected, actual, tol)                                        \
  do {                                                                                     \
    const double e_ = (expected), a_ = (actual);                                           \
    if (!(std::fabs(e_ - a_) <= (tol))) {                                                  \
      ++rpgo_failures_;                                                                    \
      std::fprintf(stderr, "%s:%d: expected %.12g got %.12g (tol %g)\n", __FILE__, __LINE__, e_, a_, \
                   (double)(tol));                                                         \
    }                                                                                      \
  } while (0)

#define RPGO_TEST_MAIN                                                                     \
  int main(int argc, char** argv) {                                                        \
    const std::string filter = argc > 1 ? argv[1] : "";                                    \
    int total = 0, failed_cases = 0;                                                       \
    for (const auto& c : rpgo_test::registry()) {                                          \
      const std::string id = std::string(c.group) + "." + c.name;                          \
      if (!filter.empty() && id.find(filter) == std::string::npos) continue;               \
      int f = 0;                                                                           \
      std::printf("[test] %s ...\n", id.c_str());                                          \
      std::fflush(stdout);                                                                 \
      try {                                                                                \
        c.fn(f);                                                                           \
      } catch (const std::exception& ex) {                                                 \
        ++f;                                                                               \
        std::fprintf(stderr, "%s: exception: %s\n", id.c_str(), ex.what());                \
      }                                                                                    \
      std::printf("[test] %s %s\n", id.c_str(), f ? "FAILED" : "ok");                     \
      std::fflush(stdout);                                                                 \
      ++total;                                                                             \
      if (f) ++failed_cases;                                                               \
    }                                                                                      \
    std::printf("[test] %d cases, %d failed\n", total, failed_cases);                    \
    return failed_cases ? 1 : 0;                                                           \
  }
