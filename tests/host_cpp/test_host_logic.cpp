// Host-side logic of the library that needs no GPU, built with AddressSanitizer
// and UndefinedBehaviorSanitizer by tests/test_sanitizers.py:
//   * the arena's first-fit allocator (pnx_admission.hpp) against a reference
//     model under random allocations and releases, including exhaustion (an
//     allocation that fails returns SIZE_MAX and changes nothing) and full
//     coalescing back to one free range;
//   * the admission decision: taped while the tape fits, checkpointed into a
//     lane short of work or with taping off, otherwise wait;
//   * the operation sequence of a served request (pnx_march.hpp): every
//     forward, recomputed and reverse step appears the right number of times
//     and every buffer index stays inside its buffer, for taped and
//     checkpointed sequences.
// Prints "ok" and exits 0, or prints the failed check and exits 1.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../../pnepma_amd/csrc/pnx_admission.hpp"
#include "../../pnepma_amd/csrc/pnx_march.hpp"

using namespace pnx;

static int failures = 0;
#define CHECK(c)                                                     \
  do {                                                               \
    if (!(c)) {                                                      \
      std::printf("check failed: %s (line %d)\n", #c, __LINE__);     \
      ++failures;                                                    \
    }                                                                \
  } while (0)

static void test_freelist() {
  std::mt19937_64 rng(7);
  const size_t size = 1 << 24;
  FreeList fl(size);
  struct Block { size_t off, bytes; };
  std::vector<Block> live;
  std::vector<char> owner(size / kAlign, 0);  // reference model: which 256-byte units are taken
  for (int it = 0; it < 20000; ++it) {
    const bool do_alloc = live.empty() || (rng() % 3 != 0);
    if (do_alloc) {
      const size_t bytes = 1 + rng() % (size / 16);
      const size_t used_before = fl.used();
      const size_t off = fl.alloc(bytes);
      if (off == SIZE_MAX) {
        CHECK(fl.used() == used_before);  // a failed allocation changes nothing
        // the model agrees that no free run is long enough
        size_t need = align_up(bytes) / kAlign, run = 0, best = 0;
        for (char o : owner) { run = o ? 0 : run + 1; if (run > best) best = run; }
        CHECK(best < need);
        continue;
      }
      CHECK(off % kAlign == 0 && off + align_up(bytes) <= size);
      for (size_t u = off / kAlign; u < (off + align_up(bytes)) / kAlign; ++u) {
        CHECK(!owner[u]);  // no overlap with a live block
        owner[u] = 1;
      }
      live.push_back({off, bytes});
    } else {
      const size_t k = rng() % live.size();
      const Block b = live[k];
      live.erase(live.begin() + (long)k);
      fl.release(b.off, b.bytes);
      for (size_t u = b.off / kAlign; u < (b.off + align_up(b.bytes)) / kAlign; ++u) owner[u] = 0;
    }
    size_t used = 0;
    for (const Block& b : live) used += align_up(b.bytes);
    CHECK(fl.used() == used);
  }
  for (const Block& b : live) fl.release(b.off, b.bytes);
  CHECK(fl.used() == 0);
  CHECK(fl.alloc(size) == 0);  // fully coalesced: the whole range is one block again
  CHECK(fl.alloc(1) == SIZE_MAX);
}

static void test_admission() {
  CHECK(admission(100, 1000, 900, 99, 32) == kAdmitTaped);
  CHECK(admission(100, 1000, 901, 0, 32) == kAdmitCheckpointed);   // lane short of work
  CHECK(admission(100, 1000, 901, 32, 32) == kAdmitWait);          // lane busy: wait for memory
  CHECK(admission(0, 0, 1, 1000, 32) == kAdmitCheckpointed);       // taping off: never wait
}

static void test_march(int n, int K) {
  const int Kc = K == 0 ? n : (K < n ? K : n), nck = K == 0 ? 0 : (n + Kc - 1) / Kc;
  const int ops = march_ops(n, nck);
  std::vector<int> fwd(n + 1, 0), rec(n + 1, 0), rev(n + 1, 0);
  int final_ops = 0;
  MarchPos pos = march_start();
  for (int t = 0; t < ops; ++t) {
    const MarchOp o = march_next(pos, n, Kc, nck);
    CHECK(o.kind >= kOpFwd && o.kind <= kOpFinal);
    CHECK(o.step >= 0 && o.step <= n);
    auto in_range = [&](int buf, int idx) {
      if (buf == kBufState) return idx == 0;
      if (buf == kBufCkpt) return idx >= 0 && idx < (nck > 0 ? nck : 1);
      return idx >= 0 && idx < Kc;
    };
    CHECK(in_range(o.src_buf, o.src_idx));
    if (o.kind == kOpFwd || o.kind == kOpRecompute) {
      CHECK(in_range(o.dst_buf, o.dst_idx));
      if (o.fuse) CHECK(in_range(o.dst2_buf, o.dst2_idx));
      if (o.ckpt_copy >= 0) CHECK(o.ckpt_copy < nck);
    }
    if (o.kind == kOpFwd) fwd[o.step]++;
    if (o.kind == kOpRecompute) rec[o.step]++;
    if (o.kind == kOpReverse) rev[o.step]++;
    if (o.kind == kOpFinal) final_ops++;
  }
  for (int k = 1; k <= n; ++k) {
    CHECK(fwd[k] == 1);
    CHECK(rec[k] == (nck == 0 ? 0 : 1));
    CHECK(rev[k] == 1);
  }
  CHECK(final_ops == 1);
}

int main() {
  test_freelist();
  test_admission();
  for (int n = 1; n <= 70; ++n)
    for (int K = 0; K <= n + 3; ++K) test_march(n, K);
  if (failures) {
    std::printf("%d checks failed\n", failures);
    return 1;
  }
  std::printf("ok\n");
  return 0;
}
