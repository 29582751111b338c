// dropin_driver.cpp -- exercises the C++ drop-in headers the way reference
// code does (WT() as called by wavelet/levelWT/WTTime.C; sdsl-shaped
// rank_support_v / select_support_mcl / wt_int as in rank/, select/, wt_int/
// WTTime.C) and writes every answer to binary files that the GPU test compares
// with the oracle.  Inputs come from the library's seeded generators, which the
// test calls with the same seeds.  Build variants: byte symbols, -DINT,
// -DINT -DLONG.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include "WT.h"
#include "rank_support_v_par.hpp"
#include "select_support_mcl_par.hpp"
#include "wt_int_par.hpp"
#include "wt_huff_par.hpp"
#include "fm_index_par.hpp"

static void dump(const std::string& path, const void* p, size_t bytes) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f || std::fwrite(p, 1, bytes, f) != bytes) { std::perror(path.c_str()); std::exit(2); }
  std::fclose(f);
}

int main(int argc, char** argv) {
  if (argc < 2) { std::fprintf(stderr, "usage: %s outdir\n", argv[0]); return 2; }
  const std::string out = argv[1];
  // ---- WT(s, n, sigma)
  {
    const uint64_t n = 100003, sigma = 1000;
    std::vector<symbol> s(n);
    wtg::check(wtg_gen_uniform(11, n, sizeof(symbol) == 1 ? 256 : sigma, (int)sizeof(symbol), s.data()));
    const uint64_t sg = sizeof(symbol) == 1 ? 256 : sigma;
    pair<WTnode*, long*> R = WT(s.data(), (uintT)n, (uintT)sg);
    int L = 0;
    for (uint64_t b = sg - 1; b > 0; b >>= 1) L++;
    dump(out + "/wt_words.bin", R.second, (n * L + 63) / 64 * 8);
    std::vector<uint64_t> nd(2 * sg * 5);
    for (uint64_t v = 0; v < 2 * sg; v++) {
      const WTnode& x = R.first[v];
      uint64_t f[5] = {(uint64_t)x.bitmapPtr, (uint64_t)x.length, (uint64_t)x.parent,
                       (uint64_t)x.leftChild, (uint64_t)x.rightChild};
      for (int k = 2; k < 5; k++) if (f[k] == (uint64_t)UINT_T_MAX) f[k] = UINT64_MAX;
      std::memcpy(&nd[v * 5], f, 40);
    }
    dump(out + "/wt_nodes.bin", nd.data(), nd.size() * 8);
    free(R.first);
    free(R.second);
  }
  // ---- rank_support_v<1,1>, select_support_mcl<1,1>, select_support_mcl<0,1>
  {
    const uint64_t nbits = 1000003, nw = (nbits + 63) / 64;
    std::vector<uint64_t> w(nw);
    wtg::check(wtg_gen_uniform(12, nw, UINT64_MAX, 8, w.data()));
    w[nw - 1] &= (1ull << (nbits & 63)) - 1;
    sdsl::bit_vector bv(nbits);
    std::memcpy(bv.data(), w.data(), nw * 8);
    sdsl::rank_support_v<1, 1> rs(&bv);
    std::vector<uint64_t> r;
    for (uint64_t i = 0; i <= nbits; i += 7) r.push_back(rs(i));
    r.push_back(rs(nbits));
    dump(out + "/rank.bin", r.data(), r.size() * 8);
    sdsl::select_support_mcl<1, 1> s1(&bv);
    sdsl::select_support_mcl<0, 1> s0(&bv);
    std::vector<uint64_t> a, b;
    for (uint64_t i = 1; i <= s1.arg_cnt(); i += 5) a.push_back(s1(i));
    for (uint64_t i = 1; i <= s0.arg_cnt(); i += 5) b.push_back(s0.select(i));
    dump(out + "/sel1.bin", a.data(), a.size() * 8);
    dump(out + "/sel0.bin", b.data(), b.size() * 8);
  }
  // ---- two-bit patterns: rank_support_v<10,2> / <01,2>, select_support_mcl<10,2> / <01,2>
  {
    const uint64_t nbits = 200003, nw = (nbits + 63) / 64;
    std::vector<uint64_t> w(nw);
    wtg::check(wtg_gen_uniform(14, nw, UINT64_MAX, 8, w.data()));
    w[nw - 1] &= (1ull << (nbits & 63)) - 1;
    w[nw - 1] |= 1ull << ((nbits - 1) & 63);  // last bit 1: a `10` in the padding
    sdsl::bit_vector bv(nbits);
    std::memcpy(bv.data(), w.data(), nw * 8);
    sdsl::rank_support_v<10, 2> r10(&bv);
    sdsl::rank_support_v<01, 2> r01(&bv);
    std::vector<uint64_t> r;
    for (uint64_t i = 0; i <= nbits; i += 3) { r.push_back(r10(i)); r.push_back(r01.rank(i)); }
    r.push_back(r10(nbits));
    r.push_back(r01(nbits));
    // the same ranks answered on the device (rank_batch over the argument words)
    {
      std::vector<uint64_t> idx;
      for (uint64_t i = 0; i <= nbits; i += 3) idx.push_back(i);
      idx.push_back(nbits);
      wtg::dbuf dw(nw * 8), di(idx.size() * 8), d10(idx.size() * 8), d01(idx.size() * 8);
      dw.upload(w.data(), nw * 8);
      di.upload(idx.data(), idx.size() * 8);
      r10.rank_batch(dw.as<uint64_t>(), di.as<uint64_t>(), idx.size(), d10.as<uint64_t>());
      r01.rank_batch(dw.as<uint64_t>(), di.as<uint64_t>(), idx.size(), d01.as<uint64_t>());
      std::vector<uint64_t> a10(idx.size()), a01(idx.size());
      d10.download(a10.data(), idx.size() * 8);
      d01.download(a01.data(), idx.size() * 8);
      for (size_t k = 0; k < idx.size(); k++)
        if (a10[k] != r[2 * k] || a01[k] != r[2 * k + 1]) {
          std::fprintf(stderr, "rank_batch pattern mismatch at %zu\n", k);
          return 3;
        }
    }
    dump(out + "/rank_pat.bin", r.data(), r.size() * 8);
    dump(out + "/rank10_dir.bin", r10.basic_block().data(), r10.basic_block().size() * 8);
    sdsl::select_support_mcl<10, 2> s10(&bv);
    sdsl::select_support_mcl<01, 2> s01(&bv);
    std::vector<uint64_t> a, b;
    a.push_back(s10.arg_cnt());
    b.push_back(s01.arg_cnt());
    for (uint64_t i = 1; i <= s10.arg_cnt(); i += 2) a.push_back(s10(i));
    for (uint64_t i = 1; i <= s01.arg_cnt(); i += 2) b.push_back(s01.select(i));
    dump(out + "/sel10.bin", a.data(), a.size() * 8);
    dump(out + "/sel01.bin", b.data(), b.size() * 8);
  }
  // ---- wt_int<>
  {
    const uint64_t n = 50000;
    std::vector<uint32_t> s(n);
    wtg::check(wtg_gen_uniform(13, n, 5000, 4, s.data()));
    sdsl::int_vector<32> iv(n);
    for (uint64_t i = 0; i < n; i++) iv[i] = s[i];
    sdsl::wt_int<> wt(iv, n);
    std::vector<uint64_t> res;
    res.push_back(wt.max_level);
    res.push_back(wt.sigma);
    for (uint64_t i = 0; i < n; i += 97) {
      const uint64_t c = s[(i * 31) % n];
      res.push_back(wt[i]);
      res.push_back(wt.rank(i, c));
      auto is = wt.inverse_select(i);
      res.push_back(is.first);
      res.push_back(is.second);
      uint64_t sel;
      try { sel = wt.select(1 + (i % 7), c); } catch (const std::logic_error&) { sel = UINT64_MAX; }
      res.push_back(sel);
    }
    dump(out + "/wtint.bin", res.data(), res.size() * 8);
    // lex_count / lex_smaller_count / range_search_2d (count and reported points)
    std::vector<uint64_t> lx;
    for (uint64_t k = 0; k < 200; k++) {
      const uint64_t i = (k * 7919) % n, j = i + (k * 104729) % (n - i + 1);
      const uint64_t c = (k % 9 == 0) ? 8192 + k : s[(k * 31) % n] + (k % 3) - 1;
      auto t = wt.lex_count(i, j, c);
      auto u = wt.lex_smaller_count(j, c);
      lx.push_back(i); lx.push_back(j); lx.push_back(c);
      lx.push_back(std::get<0>(t)); lx.push_back(std::get<1>(t)); lx.push_back(std::get<2>(t));
      lx.push_back(std::get<0>(u)); lx.push_back(std::get<1>(u));
    }
    dump(out + "/wtint_lex.bin", lx.data(), lx.size() * 8);
    std::vector<uint64_t> rs;
    for (uint64_t k = 0; k < 40; k++) {
      const uint64_t lb = (k * 7919) % n, rb = std::min<uint64_t>(n - 1, lb + (k * 3001) % 20000);
      const uint64_t vlb = (k * 613) % 5000, vrb = vlb + (k * 97) % 3000;
      auto r = wt.range_search_2d(lb, rb, vlb, vrb, k % 4 == 0);
      rs.push_back(lb); rs.push_back(rb); rs.push_back(vlb); rs.push_back(vrb);
      rs.push_back(r.first); rs.push_back(r.second.size());
      for (auto& pt : r.second) { rs.push_back(pt.first); rs.push_back(pt.second); }
    }
    dump(out + "/wtint_range.bin", rs.data(), rs.size() * 8);
    // interval_symbols: per query (i, j, k, then k triples (c, rank_c_i, rank_c_j))
    std::vector<uint64_t> is;
    std::vector<uint64_t> cs(wt.sigma + 1), ri(wt.sigma + 1), rj(wt.sigma + 1);
    for (uint64_t q = 0; q < 60; q++) {
      const uint64_t i = (q * 7919) % n, j = std::min<uint64_t>(n, i + (q % 5 == 0 ? q % 3 : (q * 131) % 4000));
      uint64_t k = 0;
      wt.interval_symbols(i, j, k, cs, ri, rj);
      is.push_back(i); is.push_back(j); is.push_back(k);
      for (uint64_t t = 0; t < k; t++) { is.push_back(cs[t]); is.push_back(ri[t]); is.push_back(rj[t]); }
    }
    dump(out + "/wtint_isym.bin", is.data(), is.size() * 8);
    // node interface: leaves reached by expand from the root (sym, size), then
    // ranges followed down the path of c (the final size is the count of c)
    std::vector<uint64_t> nd;
    std::vector<sdsl::wt_int<>::node_type> stack{wt.root()};
    while (!stack.empty()) {
      const auto v = stack.back();
      stack.pop_back();
      if (wt.empty(v)) continue;
      if (wt.is_leaf(v)) { nd.push_back(wt.sym(v)); nd.push_back(v.size); continue; }
      const auto ch = wt.expand(v);
      stack.push_back(ch.second);
      stack.push_back(ch.first);
    }
    dump(out + "/wtint_leaves.bin", nd.data(), nd.size() * 8);
    std::vector<uint64_t> ex;
    for (uint64_t q = 0; q < 40; q++) {
      const uint64_t i = (q * 7919) % n, j = std::min<uint64_t>(n - 1, i + (q * 977) % 9000);
      const uint64_t c = s[(q * 31) % n];
      auto v = wt.root();
      sdsl::range_type r{{i, j}};
      sdsl::range_vec_type rv{r, sdsl::range_type{{0, n - 1}}};
      while (!wt.is_leaf(v)) {
        const int bit = (c >> (wt.max_level - 1 - v.level)) & 1;
        const auto rr = wt.expand(v, r);
        const auto rrv = wt.expand(v, rv);
        const auto ch = wt.expand(v);
        r = bit ? rr.second : rr.first;
        rv = bit ? rrv.second : rrv.first;
        v = bit ? ch.second : ch.first;
      }
      ex.push_back(i); ex.push_back(j); ex.push_back(c);
      ex.push_back(r[1] + 1 - r[0]); ex.push_back(rv[0][1] + 1 - rv[0][0]); ex.push_back(rv[1][1] + 1 - rv[1][0]);
      ex.push_back(v.size); ex.push_back(wt.sym(v));
    }
    dump(out + "/wtint_expand.bin", ex.data(), ex.size() * 8);
    // iterators (wt_int_par.hpp:877-886)
    std::vector<uint64_t> it(wt.begin(), wt.end());
    it.push_back((uint64_t)(wt.end() - wt.begin()));
    dump(out + "/wtint_iter.bin", it.data(), it.size() * 8);
  }
  // ---- FMIndex() + wt_huff<> over the BWT (fm-index/par/WT.C:33-45)
  {
    const uint64_t n = 200003;
    std::vector<unsigned char> t(n);
    wtg::check(wtg_gen_dna(14, n, t.data()));
    sdsl::int_vector<8> bwt(n);
    FMIndex(t.data(), (long)n, bwt);
    std::vector<unsigned char> b(n);
    for (uint64_t i = 0; i < n; i++) b[i] = (unsigned char)bwt[i];
    dump(out + "/fm_bwt.bin", b.data(), n);
    sdsl::wt_huff<> wt(bwt, n);
    const unsigned char cs[6] = {'A', 'C', 'G', 'T', 'N', 0};
    std::vector<uint64_t> res;
    res.push_back(wt.sigma);
    for (uint64_t i = 0; i < n; i += 101) {
      const unsigned char c = cs[(i / 101) % 6];
      res.push_back(wt[i]);
      res.push_back(wt.rank(i, c));
      auto is = wt.inverse_select(i);
      res.push_back(is.first);
      res.push_back(is.second);
      uint64_t sel;
      try { sel = wt.select(1 + (i % 5000), c); } catch (const std::logic_error&) { sel = UINT64_MAX; }
      res.push_back(sel);
    }
    dump(out + "/huff.bin", res.data(), res.size() * 8);
    // interval_symbols over the wt_huff: (i, j, k, then k triples (c, rank_c_i, rank_c_j))
    std::vector<uint64_t> hs;
    std::vector<uint8_t> hcs(wt.sigma + 1);
    std::vector<uint64_t> hri(wt.sigma + 1), hrj(wt.sigma + 1);
    for (uint64_t q = 0; q < 60; q++) {
      const uint64_t i = (q * 7919) % n, j = std::min<uint64_t>(n, i + (q % 4 == 0 ? q % 4 + (q / 4) % 3 : (q * 131) % 5000));
      uint64_t k = 0;
      wt.interval_symbols(i, j, k, hcs, hri, hrj);
      hs.push_back(i); hs.push_back(j); hs.push_back(k);
      for (uint64_t t = 0; t < k; t++) { hs.push_back(hcs[t]); hs.push_back(hri[t]); hs.push_back(hrj[t]); }
    }
    dump(out + "/huff_isym.bin", hs.data(), hs.size() * 8);
    // iterators (wt_pc_par.hpp:740-746)
    std::vector<unsigned char> hit(wt.begin(), wt.end());
    dump(out + "/huff_iter.bin", hit.data(), hit.size());
    // node interface (wt_pc_par.hpp:778-909): leaves by expand from the root with
    // the root range [0, n-1] mapped down (sym, range size); then, per symbol
    // c, the range [i, j] followed down c's path (path length, code, final
    // range size, leaf sym); symbol_gte / symbol_lte for every byte
    std::vector<uint64_t> hl;
    std::vector<std::pair<sdsl::wt_huff<>::node_type, sdsl::range_type>> st{{wt.root(), sdsl::range_type{{0, n - 1}}}};
    while (!st.empty()) {
      const auto e = st.back();
      st.pop_back();
      const uint64_t sz = e.second[1] + 1 - e.second[0];
      if (!wt.empty(e.first) || sz == 0) continue;
      if (wt.is_leaf(e.first)) { hl.push_back(wt.sym(e.first)); hl.push_back(sz); continue; }
      const auto ch = wt.expand(e.first);
      const auto rr = wt.expand(e.first, e.second);
      st.push_back({ch.second, rr.second});
      st.push_back({ch.first, rr.first});
    }
    dump(out + "/huff_leaves.bin", hl.data(), hl.size() * 8);
    std::vector<uint64_t> hp;
    for (int c = 0; c < 256; c++) {
      const auto g = wt.symbol_gte((unsigned char)c), l = wt.symbol_lte((unsigned char)c);
      hp.push_back(g.first); hp.push_back(g.second); hp.push_back(l.first); hp.push_back(l.second);
      if (!g.first || g.second != c) continue;
      const auto pl = wt.path((unsigned char)c);
      const uint64_t i = (c * 7919) % n, j = std::min<uint64_t>(n - 1, i + 20000);
      sdsl::range_type r{{i, j}};
      sdsl::range_vec_type rv{r, sdsl::range_type{{0, n - 1}}};
      auto v = wt.root();
      for (uint64_t d = 0; d < pl.first; d++) {
        const int bit = (pl.second >> (pl.first - 1 - d)) & 1;
        const auto rr = wt.expand(v, r);
        const auto rrv = wt.expand(v, rv);
        r = bit ? rr.second : rr.first;
        rv = bit ? rrv.second : rrv.first;
        v = bit ? wt.expand(v).second : wt.expand(v).first;
      }
      hp.push_back(c); hp.push_back(pl.first); hp.push_back(pl.second); hp.push_back(i); hp.push_back(j);
      hp.push_back(r[1] + 1 - r[0]); hp.push_back(rv[1][1] + 1 - rv[1][0]); hp.push_back(wt.is_leaf(v));
      hp.push_back(wt.sym(v));
    }
    dump(out + "/huff_nodes.bin", hp.data(), hp.size() * 8);
    // host backward-search counts through the fm_index wrapper
    wtg::fm_index idx(t.data(), n);
    const uint32_t m = 12;
    const uint64_t np = 300;
    std::vector<unsigned char> pats(np * m);
    wtg::check(wtg_gen_patterns(15, t.data(), n, m, np, pats.data()));
    std::vector<uint64_t> cnt(np);
    for (uint64_t q = 0; q < np; q++) cnt[q] = idx.count(pats.data() + q * m, m);
    dump(out + "/fm_count.bin", cnt.data(), np * 8);
  }
  std::printf("dropin ok\n");
  return 0;
}
