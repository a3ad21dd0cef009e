// A header-only TEST / EXPECT_* harness with the gtest spellings the
// reference's tests use (gtest itself is not installed here).  Each binary
// runs every registered test (or those whose "Suite.Name" contains argv[1])
// and exits non-zero if any expectation failed.
#pragma once

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <limits>
#include <sstream>
#include <string>
#include <vector>

#include <sys/mman.h>
#include <unistd.h>

namespace mt {

struct Case {
    std::string name;
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

struct Registrar {
    Registrar(const char* suite, const char* name, std::function<void()> fn) {
        registry().push_back({std::string(suite) + "." + name, std::move(fn)});
    }
};

// Collects a streamed message; records the failure when destroyed.
class Failure {
public:
    Failure(const char* file, int line, std::string what) {
        os_ << file << ":" << line << ": " << what;
    }
    ~Failure() {
        std::fprintf(stderr, "  FAILED %s\n", os_.str().c_str());
        ++failures();
    }
    template <class V>
    Failure& operator<<(const V& v) {
        os_ << " " << v;
        return *this;
    }

private:
    std::ostringstream os_;
};

template <class A, class B>
std::string pair_str(const char* op, const A& a, const B& b) {
    std::ostringstream os;
    os.precision(17);
    os << "(" << a << " " << op << " " << b << ")";
    return os.str();
}

inline bool float_eq(float a, float b) {
    if (a == b) return true;
    if (std::isnan(a) || std::isnan(b)) return false;
    int32_t ia, ib;
    std::memcpy(&ia, &a, 4);
    std::memcpy(&ib, &b, 4);
    if ((ia < 0) != (ib < 0)) return false;
    return std::abs(ia - ib) <= 4;  // gtest: within 4 ULPs
}
inline bool double_eq(double a, double b) {
    if (a == b) return true;
    if (std::isnan(a) || std::isnan(b)) return false;
    int64_t ia, ib;
    std::memcpy(&ia, &a, 8);
    std::memcpy(&ib, &b, 8);
    if ((ia < 0) != (ib < 0)) return false;
    return std::llabs(ia - ib) <= 4;
}

// Captures what a callable writes to stdout.
inline std::string capture_stdout(const std::function<void()>& fn) {
    std::fflush(stdout);
    const int fd = memfd_create("mt_stdout", 0);  // anonymous in-memory file
    const int saved = dup(1);
    dup2(fd, 1);
    fn();
    std::fflush(stdout);
    dup2(saved, 1);
    close(saved);
    std::string out;
    lseek(fd, 0, SEEK_SET);
    char buf[4096];
    ssize_t n;
    while ((n = read(fd, buf, sizeof(buf))) > 0) out.append(buf, static_cast<size_t>(n));
    close(fd);
    return out;
}

inline int run_all(int argc, char** argv) {
    const char* filter = argc > 1 ? argv[1] : "";
    int ran = 0, failed = 0;
    for (auto& c : registry()) {
        if (*filter && c.name.find(filter) == std::string::npos) continue;
        const int before = failures();
        std::printf("[ RUN  ] %s\n", c.name.c_str());
        std::fflush(stdout);
        try {
            c.fn();
        } catch (const std::exception& e) {
            std::fprintf(stderr, "  FAILED uncaught exception: %s\n", e.what());
            ++failures();
        }
        ++ran;
        const bool ok = failures() == before;
        failed += ok ? 0 : 1;
        std::printf("[ %s ] %s\n", ok ? " OK " : "FAIL", c.name.c_str());
        std::fflush(stdout);
    }
    std::printf("%d tests, %d failed\n", ran, failed);
    return (failed == 0 && ran > 0) ? 0 : 1;
}

}  // namespace mt

#define MT_CAT2(a, b) a##b
#define MT_CAT(a, b) MT_CAT2(a, b)
#define TEST(suite, name)                                                               \
    static void MT_CAT(suite, MT_CAT(_, name))();                                       \
    static ::mt::Registrar MT_CAT(reg_, MT_CAT(suite, MT_CAT(_, name)))(#suite, #name,  \
                                                                      &MT_CAT(suite, MT_CAT(_, name))); \
    static void MT_CAT(suite, MT_CAT(_, name))()

#define MT_CHECK(cond, what) \
    if (cond) {              \
    } else                   \
        ::mt::Failure(__FILE__, __LINE__, what)
#define MT_ASSERT(cond, what)                            \
    if (cond) {                                          \
    } else                                               \
        return (void)(::mt::Failure(__FILE__, __LINE__, what))

#define EXPECT_TRUE(c) MT_CHECK((c), "EXPECT_TRUE(" #c ")")
#define EXPECT_FALSE(c) MT_CHECK(!(c), "EXPECT_FALSE(" #c ")")
#define EXPECT_EQ(a, b) MT_CHECK((a) == (b), "EXPECT_EQ" + ::mt::pair_str("==", (a), (b)))
#define EXPECT_NE(a, b) MT_CHECK((a) != (b), "EXPECT_NE" + ::mt::pair_str("!=", (a), (b)))
#define EXPECT_LT(a, b) MT_CHECK((a) < (b), "EXPECT_LT" + ::mt::pair_str("<", (a), (b)))
#define EXPECT_LE(a, b) MT_CHECK((a) <= (b), "EXPECT_LE" + ::mt::pair_str("<=", (a), (b)))
#define EXPECT_GT(a, b) MT_CHECK((a) > (b), "EXPECT_GT" + ::mt::pair_str(">", (a), (b)))
#define EXPECT_GE(a, b) MT_CHECK((a) >= (b), "EXPECT_GE" + ::mt::pair_str(">=", (a), (b)))
#define EXPECT_NEAR(a, b, tol) \
    MT_CHECK(std::abs(double(a) - double(b)) <= double(tol), "EXPECT_NEAR" + ::mt::pair_str("~", (a), (b)))
#define EXPECT_FLOAT_EQ(a, b) \
    MT_CHECK(::mt::float_eq(float(a), float(b)), "EXPECT_FLOAT_EQ" + ::mt::pair_str("==", (a), (b)))
#define EXPECT_DOUBLE_EQ(a, b) \
    MT_CHECK(::mt::double_eq(double(a), double(b)), "EXPECT_DOUBLE_EQ" + ::mt::pair_str("==", (a), (b)))
#define EXPECT_THROW(stmt, exc)                                  \
    do {                                                         \
        bool _thrown = false;                                    \
        try {                                                    \
            stmt;                                                \
        } catch (const exc&) {                                   \
            _thrown = true;                                      \
        } catch (...) {                                          \
        }                                                        \
        MT_CHECK(_thrown, "EXPECT_THROW(" #stmt ", " #exc ")");  \
    } while (0)
#define EXPECT_NO_THROW(stmt)                                    \
    do {                                                         \
        bool _thrown = false;                                    \
        try {                                                    \
            stmt;                                                \
        } catch (...) {                                          \
            _thrown = true;                                      \
        }                                                        \
        MT_CHECK(!_thrown, "EXPECT_NO_THROW(" #stmt ")");        \
    } while (0)
#define ASSERT_NE(a, b) MT_ASSERT((a) != (b), "ASSERT_NE" + ::mt::pair_str("!=", (a), (b)))
#define ASSERT_GT(a, b) MT_ASSERT((a) > (b), "ASSERT_GT" + ::mt::pair_str(">", (a), (b)))
#define ASSERT_TRUE(c) MT_ASSERT((c), "ASSERT_TRUE(" #c ")")

#define MT_MAIN() \
    int main(int argc, char** argv) { return ::mt::run_all(argc, argv); }
