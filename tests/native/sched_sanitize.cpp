// Sanitizer driver of the native scheduler (schedule.cpp) and the C-ABI
// error state (capi.cpp), built by tests/test_native_sanitize.py with
// -fsanitize=address,undefined.  Reads one operator's directed faces from a
// binary file, schedules every direction, checks the exported arrays, then
// repeats with operator new failing after a countdown (inside the OpenMP
// region) and expects OBTE_ENOMEM with nothing leaked.
#include <atomic>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <new>
#include <vector>

#include "openbte_amd.h"

static std::atomic<long> g_fail_after{-1};  // < 0: never fail

void* operator new(std::size_t n) {
  long k = g_fail_after.load();
  if (k >= 0 && g_fail_after.fetch_sub(1) == 0) throw std::bad_alloc();
  if (void* p = std::malloc(n ? n : 1)) return p;
  throw std::bad_alloc();
}
void operator delete(void* p) noexcept { std::free(p); }
void operator delete(void* p, std::size_t) noexcept { std::free(p); }

#define CHECK(c)                                                   \
  do {                                                             \
    if (!(c)) {                                                    \
      std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
      return 1;                                                    \
    }                                                              \
  } while (0)

template <class T>
static bool read_vec(FILE* f, std::vector<T>& v, int64_t n) {
  v.resize(n);
  return std::fread(v.data(), sizeof(T), n, f) == (size_t)n;
}

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  FILE* f = std::fopen(argv[1], "rb");
  if (!f) return 2;
  int64_t hdr[4];  // n, faces, dim, dirs
  if (std::fread(hdr, sizeof(int64_t), 4, f) != 4) return 2;
  const int32_t n = (int32_t)hdr[0], dim = (int32_t)hdr[2], L = (int32_t)hdr[3];
  const int64_t F = hdr[1];
  std::vector<int32_t> im, jm;
  std::vector<double> k, dirs;
  std::vector<uint8_t> cut;
  if (!read_vec(f, im, F) || !read_vec(f, jm, F) || !read_vec(f, k, F * dim) || !read_vec(f, cut, F) ||
      !read_vec(f, dirs, (int64_t)L * dim))
    return 2;
  std::fclose(f);

  void* h = nullptr;
  CHECK(obte_sched_create(n, F, dim, im.data(), jm.data(), k.data(), cut.data(), L, dirs.data(), 192, &h) == 0);
  std::vector<int64_t> sizes(3 * L);
  CHECK(obte_sched_sizes(h, sizes.data()) == 0);
  int64_t lv = 0, ent = 0, cuts = 0;
  for (int l = 0; l < L; ++l) {
    lv += sizes[3 * l] + 1;
    ent += sizes[3 * l + 1];
    cuts += sizes[3 * l + 2];
  }
  std::vector<int32_t> order((int64_t)L * n), pos((int64_t)L * n), lvl(lv), inp((int64_t)L * (n + 1)), src(ent + 1),
      face(ent + 1), cex(cuts + 1), cfc(cuts + 1);
  CHECK(obte_sched_export(h, order.data(), pos.data(), lvl.data(), inp.data(), src.data(), face.data(), cex.data(),
                          cfc.data()) == 0);
  const int kin = 3;
  std::vector<int32_t> nb((int64_t)L * kin * n), nf((int64_t)L * kin * n);
  CHECK(obte_sched_export_slots(h, kin, nb.data(), nf.data()) == 0);
  CHECK(obte_sched_export_slots(h, 0, nb.data(), nf.data()) == OBTE_EINVAL);
  // every direction: a permutation, its inverse, upstream positions earlier
  int64_t e0 = 0;
  for (int l = 0; l < L; ++l) {
    std::vector<int> seen(n, 0);
    for (int p = 0; p < n; ++p) {
      const int c = order[(int64_t)l * n + p];
      CHECK(c >= 0 && c < n && !seen[c]);
      seen[c] = 1;
      CHECK(pos[(int64_t)l * n + c] == p);
      const int a = inp[(int64_t)l * (n + 1) + p], b = inp[(int64_t)l * (n + 1) + p + 1];
      CHECK(b - a <= kin);
      for (int e = a; e < b; ++e) {
        const int s = src[e0 + e];
        if (s >= 0) CHECK(s < p);
        CHECK(nb[((int64_t)l * kin + (e - a)) * n + p] == (s >= 0 ? s : n + ~s));
      }
    }
    e0 += sizes[3 * l + 1];
  }
  CHECK(obte_sched_destroy(h) == 0);

  // allocation failure inside the parallel region: reported, nothing leaked
  for (long after : {0L, 5L, 200L, 5000L}) {
    void* h2 = nullptr;
    g_fail_after.store(after);
    const int rc = obte_sched_create(n, F, dim, im.data(), jm.data(), k.data(), cut.data(), L, dirs.data(), 192, &h2);
    g_fail_after.store(-1);
    if (rc == 0) {
      CHECK(h2 != nullptr);
      obte_sched_destroy(h2);
    } else {
      CHECK(rc == OBTE_ENOMEM && h2 == nullptr);
      std::printf("forced failure after %ld allocations: %s\n", after, obte_last_error());
    }
  }
  std::printf("ok %d directions\n", L);
  return 0;
}
