// serialize_driver.cpp -- serialize / load round trips of the C++ drop-ins
// (rank_support_v<1> and <0>, select_support_mcl<1> and <0>, wt_int<>,
// wt_huff<>) in the reference's stream layouts.  Every structure is built on
// the GPU, written to <outdir>/<name>.bin, loaded into a fresh object and
// checked here: the loaded object answers every query like the original (its
// device copy too), and writing it again gives the same bytes.  The GPU test
// then decodes the written images independently and compares their fields
// with the oracle.  Exit status 0 = all checks passed.
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>
#include "rank_support_v_par.hpp"
#include "select_support_mcl_par.hpp"
#include "wt_huff_par.hpp"
#include "wt_int_par.hpp"

static int g_fail = 0;
#define CHECK(cond, what)                                                   \
  do {                                                                      \
    if (!(cond)) { std::fprintf(stderr, "FAIL: %s\n", what); g_fail++; }    \
  } while (0)

template <class T>
static std::string image(const T& x) {
  std::ostringstream o;
  x.serialize(o);
  return o.str();
}
static void save(const std::string& path, const std::string& bytes) {
  std::ofstream f(path, std::ios::binary);
  f.write(bytes.data(), (std::streamsize)bytes.size());
}

template <class R, class S1, class S0>
static void supports(const std::string& out, const std::string& tag, const sdsl::bit_vector& bv) {
  R r1(&bv);
  sdsl::rank_support_v<0, 1> r0(&bv);
  S1 s1(&bv);
  S0 s0(&bv);
  const std::string ir1 = image(r1), ir0 = image(r0), is1 = image(s1), is0 = image(s0);
  save(out + "/rank1_" + tag + ".bin", ir1);
  save(out + "/rank0_" + tag + ".bin", ir0);
  save(out + "/sel1_" + tag + ".bin", is1);
  save(out + "/sel0_" + tag + ".bin", is0);
  R l1; sdsl::rank_support_v<0, 1> l0; S1 m1; S0 m0;
  { std::istringstream in(ir1); l1.load(in, &bv); }
  { std::istringstream in(ir0); l0.load(in, &bv); }
  { std::istringstream in(is1); m1.load(in, &bv); }
  { std::istringstream in(is0); m0.load(in, &bv); }
  const uint64_t n = bv.size();
  uint64_t ones = 0;
  for (uint64_t i = 0; i <= n; i++) {
    if (i % 61 == 0 || i + 70 > n) {
      CHECK(l1(i) == r1(i) && l1(i) == ones, "rank1 after load");
      CHECK(l0(i) == r0(i) && l0(i) == i - ones, "rank0 after load");
    }
    if (i < n) ones += bv[i];
  }
  for (uint64_t i = 1; i <= s1.arg_cnt(); i += 3) CHECK(m1(i) == s1(i), "select1 after load");
  for (uint64_t i = 1; i <= s0.arg_cnt(); i += 3) CHECK(m0(i) == s0(i), "select0 after load");
  CHECK(m1.arg_cnt() == s1.arg_cnt() && m0.arg_cnt() == s0.arg_cnt(), "arg counts after load");
  CHECK(image(l1) == ir1 && image(l0) == ir0 && image(m1) == is1 && image(m0) == is0, "re-serialized supports");
}

int main(int argc, char** argv) {
  if (argc < 2) { std::fprintf(stderr, "usage: %s outdir\n", argv[0]); return 2; }
  const std::string out = argv[1];
  wtg_ctx* ctx = wtg::default_context();
  // ---- supports over a dense random vector and a sparse one with long superblocks
  {
    const uint64_t nbits = 1000003, nw = (nbits + 63) / 64;
    sdsl::bit_vector bv(nbits);
    wtg::check(wtg_gen_uniform(12, nw, UINT64_MAX, 8, bv.data()));
    bv.data()[nw - 1] &= (1ull << (nbits & 63)) - 1;
    supports<sdsl::rank_support_v<1, 1>, sdsl::select_support_mcl<1, 1>, sdsl::select_support_mcl<0, 1>>(
        out, "dense", bv);
    sdsl::bit_vector sp((3ull << 20) + 5);
    for (uint64_t i = 0; i < sp.size(); i += 331) sp.set(i, true);
    supports<sdsl::rank_support_v<1, 1>, sdsl::select_support_mcl<1, 1>, sdsl::select_support_mcl<0, 1>>(
        out, "sparse", sp);
  }
  // ---- wt_int<>
  {
    const uint64_t n = 50000;
    std::vector<uint32_t> s(n);
    wtg::check(wtg_gen_uniform(13, n, 5000, 4, s.data()));
    sdsl::int_vector<32> iv(n);
    for (uint64_t i = 0; i < n; i++) iv[i] = s[i];
    sdsl::wt_int<> wt(iv, n);
    const std::string im = image(wt);
    save(out + "/wtint.bin", im);
    sdsl::wt_int<> lw;
    { std::istringstream in(im); lw.load(in); }
    CHECK(lw.size() == n && lw.sigma == wt.sigma && lw.max_level == wt.max_level, "wt_int header after load");
    for (uint64_t i = 0; i < n; i += 37) {
      const uint64_t c = s[(i * 31) % n];
      CHECK(lw[i] == s[i], "wt_int access after load");
      CHECK(lw.rank(i, c) == wt.rank(i, c), "wt_int rank after load");
      CHECK(lw.inverse_select(i) == wt.inverse_select(i), "wt_int inverse_select after load");
      CHECK(lw.select(1, s[i]) == wt.select(1, s[i]), "wt_int select after load");
    }
    CHECK(image(lw) == im, "re-serialized wt_int");
    // the loaded object's device copy answers batches like the original's
    const uint64_t nq = 4096;
    std::vector<uint64_t> qi(nq), qc(nq), a(nq), b(nq);
    for (uint64_t q = 0; q < nq; q++) { qi[q] = (q * 7919) % (n + 1); qc[q] = s[(q * 13) % n]; }
    wtg::dbuf di(nq * 8), dc(nq * 8), da(nq * 8), db(nq * 8);
    di.upload(qi.data(), nq * 8);
    dc.upload(qc.data(), nq * 8);
    wt.rank_batch(di.as<uint64_t>(), dc.as<uint64_t>(), nq, da.as<uint64_t>());
    lw.rank_batch(di.as<uint64_t>(), dc.as<uint64_t>(), nq, db.as<uint64_t>());
    da.download(a.data(), nq * 8);
    db.download(b.data(), nq * 8);
    CHECK(a == b, "wt_int rank_batch after load");
    sdsl::wt_int<> e0, e1;
    const std::string ie = image(e0);
    { std::istringstream in(ie); e1.load(in); }
    CHECK(e1.size() == 0 && image(e1) == ie, "empty wt_int round trip");
  }
  // ---- wt_huff<>
  {
    const uint64_t n = 200000;
    std::vector<uint8_t> t(n);
    wtg::check(wtg_gen_dna(14, n, t.data()));
    sdsl::wt_huff<> wt(t.data(), n);
    const std::string im = image(wt);
    save(out + "/huff.bin", im);
    sdsl::wt_huff<> lw;
    { std::istringstream in(im); lw.load(in); }
    CHECK(lw.size() == n && lw.sigma == wt.sigma, "wt_huff header after load");
    for (uint64_t i = 0; i < n; i += 41) {
      const uint8_t c = t[(i * 31) % n];
      CHECK(lw[i] == t[i], "wt_huff access after load");
      CHECK(lw.rank(i, c) == wt.rank(i, c), "wt_huff rank after load");
      CHECK(lw.select(1 + i % 5, c) == wt.select(1 + i % 5, c), "wt_huff select after load");
    }
    CHECK(image(lw) == im, "re-serialized wt_huff");
    // FM counts through the loaded object's device copy
    const uint32_t m = 8;
    const uint64_t np = 512;
    std::vector<uint8_t> pats(np * m);
    for (uint64_t q = 0; q < np; q++) std::memcpy(&pats[q * m], &t[(q * 104729) % (n - m - 1)], m);
    wtg::dbuf dp(np * m), da(np * 8), db(np * 8);
    dp.upload(pats.data(), np * m);
    wt.count_batch((const uint8_t*)dp.get(), m, np, da.as<uint64_t>());
    lw.count_batch((const uint8_t*)dp.get(), m, np, db.as<uint64_t>());
    std::vector<uint64_t> a(np), b(np);
    da.download(a.data(), np * 8);
    db.download(b.data(), np * 8);
    CHECK(a == b, "wt_huff count_batch after load");
    // a stored tree that is not the Huffman tree of its counts is rejected:
    // swap the children of the root in the image (node 0 sits right after the
    // supports; its child fields are the last 4 bytes of its 22-byte record)
    std::string bad = im;
    const uint64_t nn_at = im.size() - 256 * 2 - 256 * 8 - (uint64_t)wt.shape().n_nodes * 22 - 8;
    char* rec0 = &bad[nn_at + 8];
    std::swap_ranges(rec0 + 18, rec0 + 20, rec0 + 20);
    bool threw = false;
    try {
      sdsl::wt_huff<> x;
      std::istringstream in(bad);
      x.load(in);
    } catch (const std::runtime_error&) {
      threw = true;
    }
    CHECK(threw, "wt_huff load rejects a tree that is not the Huffman tree of its counts");
    sdsl::wt_huff<> e0, e1;
    const std::string ie = image(e0);
    { std::istringstream in(ie); e1.load(in); }
    CHECK(e1.size() == 0 && image(e1) == ie, "empty wt_huff round trip");
  }
  (void)ctx;
  if (g_fail) std::fprintf(stderr, "%d check(s) failed\n", g_fail);
  else std::printf("serialize round trips ok\n");
  return g_fail ? 1 : 0;
}
