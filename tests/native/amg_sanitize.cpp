// Sanitizer driver of the native multigrid setup (amg.cpp) and the C-ABI
// error state (capi.cpp), built by tests/test_native_sanitize.py with
// -fsanitize=address,undefined.  Reads one Fourier operator (rows, columns,
// off-diagonal weights, diagonal) from a binary file, builds the hierarchy,
// checks every exported level (monotone row pointers, columns in range,
// R = P^T entry for entry), writes the exported arrays to an output file for
// a bitwise comparison with the production library, then repeats with
// operator new failing after a countdown and expects OBTE_ENOMEM with nothing
// leaked.
#include <atomic>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <new>
#include <utility>
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

#define CHECK(c)                                                          \
  do {                                                                    \
    if (!(c)) {                                                           \
      std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
      return 1;                                                           \
    }                                                                     \
  } while (0)

template <class T>
static bool read_vec(FILE* f, std::vector<T>& v, int64_t n) {
  v.resize(n);
  return std::fread(v.data(), sizeof(T), n, f) == (size_t)n;
}

template <class T>
static void write_vec(FILE* f, const std::vector<T>& v) {
  std::fwrite(v.data(), sizeof(T), v.size(), f);
}

static bool csr_ok(const std::vector<int32_t>& ptr, const std::vector<int32_t>& col, int32_t rows, int32_t cols) {
  if ((int32_t)ptr.size() != rows + 1 || ptr[0] != 0) return false;
  for (int32_t i = 0; i < rows; ++i)
    if (ptr[i + 1] < ptr[i]) return false;
  for (int32_t k = 0; k < ptr[rows]; ++k)
    if (col[k] < 0 || col[k] >= cols) return false;
  return true;
}

int main(int argc, char** argv) {
  if (argc < 3) return 2;
  FILE* f = std::fopen(argv[1], "rb");
  if (!f) return 2;
  int64_t hdr[2];
  if (std::fread(hdr, sizeof(int64_t), 2, f) != 2) return 2;
  const int32_t n = (int32_t)hdr[0];
  const int64_t nnz = hdr[1];
  std::vector<int32_t> row_ptr, col;
  std::vector<double> w, diag;
  if (!read_vec(f, row_ptr, n + 1) || !read_vec(f, col, nnz) || !read_vec(f, w, nnz) || !read_vec(f, diag, n))
    return 2;
  std::fclose(f);

  void* h = nullptr;
  CHECK(obte_amg_create(n, row_ptr.data(), col.data(), w.data(), diag.data(), 8, 512, 2.0 / 3.0, 16, &h) ==
        OBTE_OK);
  int32_t nl = 0;
  CHECK(obte_amg_sizes(h, &nl, nullptr) == OBTE_OK && nl >= 2);
  std::vector<int64_t> sizes(4 * (size_t)nl);
  CHECK(obte_amg_sizes(h, &nl, sizes.data()) == OBTE_OK);
  FILE* out = std::fopen(argv[2], "wb");
  CHECK(out != nullptr);
  for (int32_t l = 0; l + 1 < nl; ++l) {
    const int32_t rows = (int32_t)sizes[4 * l], nn = (int32_t)sizes[4 * (l + 1)];
    std::vector<int32_t> ap(rows + 1), ac(sizes[4 * l + 1] > 0 ? sizes[4 * l + 1] : 1);
    std::vector<double> aw(ac.size()), ad(rows);
    std::vector<int32_t> pp(rows + 1), pc(sizes[4 * l + 2]), rp(nn + 1), rc(sizes[4 * l + 3]);
    std::vector<double> pv(pc.size()), rv(rc.size());
    CHECK(obte_amg_export(h, l, ap.data(), ac.data(), aw.data(), ad.data(), pp.data(), pc.data(), pv.data(),
                          rp.data(), rc.data(), rv.data()) == OBTE_OK);
    CHECK(csr_ok(ap, ac, rows, rows));
    CHECK(csr_ok(pp, pc, rows, nn));
    CHECK(csr_ok(rp, rc, nn, rows));
    CHECK(pc.size() == rc.size());
    // R = P^T entry for entry
    std::map<std::pair<int32_t, int32_t>, double> pe;
    for (int32_t i = 0; i < rows; ++i)
      for (int32_t k = pp[i]; k < pp[i + 1]; ++k) pe[{i, pc[k]}] = pv[k];
    for (int32_t I = 0; I < nn; ++I)
      for (int32_t k = rp[I]; k < rp[I + 1]; ++k) {
        auto it = pe.find({rc[k], I});
        CHECK(it != pe.end() && it->second == rv[k]);
      }
    for (int32_t i = 0; i < rows; ++i) CHECK(ad[i] > 0.0);
    write_vec(out, pp);
    write_vec(out, pc);
    write_vec(out, pv);
    if (l > 0) {
      write_vec(out, ap);
      ac.resize(ap[rows]);
      aw.resize(ap[rows]);
      write_vec(out, ac);
      write_vec(out, aw);
      write_vec(out, ad);
    }
  }
  const int64_t nc = sizes[4 * (nl - 1)];
  std::vector<double> dense((size_t)(nc * nc));
  CHECK(obte_amg_export_dense(h, dense.data()) == OBTE_OK);
  write_vec(out, dense);
  std::fclose(out);
  CHECK(obte_amg_destroy(h) == OBTE_OK);
  // invalid arguments fail with a message, not a crash
  CHECK(obte_amg_create(n, row_ptr.data(), col.data(), w.data(), diag.data(), 1, 512, 2.0 / 3.0, 16, &h) ==
        OBTE_EINVAL);
  CHECK(obte_last_error()[0] != '\0');
  std::printf("ok %d levels\n", (int)nl);

  // operator new fails after a countdown inside the build: OBTE_ENOMEM, no leak
  int failures = 0;
  for (long after : {0L, 3L, 50L, 400L}) {
    g_fail_after.store(after);
    void* h2 = nullptr;
    const int rc = obte_amg_create(n, row_ptr.data(), col.data(), w.data(), diag.data(), 8, 512, 2.0 / 3.0, 16, &h2);
    g_fail_after.store(-1);
    if (rc == OBTE_ENOMEM) {
      CHECK(h2 == nullptr);
      ++failures;
    } else {
      CHECK(rc == OBTE_OK);
      obte_amg_destroy(h2);
    }
  }
  CHECK(failures > 0);
  std::printf("forced failure %d\n", failures);
  return 0;
}
