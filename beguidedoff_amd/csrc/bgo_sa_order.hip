// Anchor occurrences in the suffix-array order of the reference's fast-mode
// index text (host code).
//
// The reference's fast mode locates an anchor in an FM-index of the genome
// and keeps the first `max_hits_per_anchor` entries of the anchor's
// suffix-array interval (csrc/fmindex_binding.cpp:164-176,
// fast/finder.py:149-159).  Its text (fast/fmindex.py:63-128,
// fast/backend.py:28-49) is the contigs' bases as symbols A C G T = 1 2 3 4
// (N and every other byte = 1, as A), a separator symbol 6 between contigs,
// and the terminal 0 that the suffix-array construction appends.  The
// occurrences it keeps are therefore the `cap` occurrences whose suffixes of
// that text are lexicographically smallest.
//
// Here the genome is the 2-bit forward array (N coded as A, as in the text).
// The selection refines the occurrences group by group on runs of equal
// symbols rather than single symbols:
//   the suffix at a cursor starts with a run (s, r) (symbol s repeated r
//   times) followed by a different symbol b.  Between two suffixes that start
//   with runs of the same s, a shorter run is smaller iff the symbol after it
//   is smaller than s, so the key (s, b > s, b > s ? -r : r) orders them,
//   and equal keys mean equal runs: the cursors then advance by r together.
// Only the group that straddles the cap is refined further.  Maximal runs of
// >= 64 equal bases (N blocks, homopolymers) are listed once per genome, so
// the run length at any cursor costs a table search or a scan of < 64 bases;
// an anchor with a hundred million occurrences inside N blocks (A^12) is
// settled by one key per occurrence and a selection.
#include <algorithm>
#include <cstring>
#include <thread>
#include <vector>

#include "bgo_common.h"

struct bgo_sa_text {
  const uint32_t* words = nullptr;  // caller-owned, outlives the handle
  int64_t total_bp = 0;
  std::vector<int64_t> off, len;    // contigs
  std::vector<int64_t> run_start, run_end;  // maximal same-code runs >= kLongRun, by start
};

namespace {

using bgo::fail;

constexpr int64_t kLongRun = 64;
constexpr uint64_t kMaxR = (uint64_t{1} << 40) - 1;

inline uint32_t code_at(const uint32_t* w, int64_t i) {
  return (w[i >> 4] >> ((i & 15) * 2)) & 3u;
}

// Maximal runs of one code of >= kLongRun bases inside one contig.  Such a
// run covers at least three whole, equal, uniform 16-base words.
void find_long_runs(bgo_sa_text* T) {
  for (size_t k = 0; k < T->off.size(); ++k) {
    const int64_t a = T->off[k], b = a + T->len[k];
    const int64_t w1 = b / 16;
    int64_t w = (a + 15) / 16;
    while (w < w1) {
      const uint32_t x = T->words[w];
      if (!(x == 0u || x == 0x55555555u || x == 0xAAAAAAAAu || x == 0xFFFFFFFFu)) {
        ++w;
        continue;
      }
      int64_t e = w + 1;
      while (e < w1 && T->words[e] == x) ++e;
      if (e - w >= 3) {
        const uint32_t c = x & 3u;
        int64_t s = w * 16, t = e * 16;
        while (s > a && code_at(T->words, s - 1) == c) --s;
        while (t < b && code_at(T->words, t) == c) ++t;
        if (t - s >= kLongRun) {
          T->run_start.push_back(s);
          T->run_end.push_back(t);
        }
      }
      w = e;
    }
  }
}

// Length of the run of equal codes starting at g, inside contig end `end`.
int64_t run_length(const bgo_sa_text* T, int64_t g, int64_t end) {
  const auto it = std::upper_bound(T->run_start.begin(), T->run_start.end(), g);
  if (it != T->run_start.begin()) {
    const size_t j = static_cast<size_t>(it - T->run_start.begin()) - 1;
    if (g < T->run_end[j]) return std::min(T->run_end[j], end) - g;
  }
  const uint32_t c = code_at(T->words, g);
  int64_t r = 1;
  while (g + r < end && code_at(T->words, g + r) == c) ++r;  // < kLongRun steps
  return r;
}

struct Cursor {
  uint32_t q;  // the occurrence (answer)
  uint32_t k;  // contig of the cursor
  int64_t g;   // cursor; == contig end: the separator (or the terminal 0)
  int64_t r;   // length of the run at the cursor (set with the key)
};

// Key of the run at the cursor (smaller key = smaller suffix).
uint64_t run_key(const bgo_sa_text* T, Cursor& c) {
  const int64_t end = T->off[c.k] + T->len[c.k];
  const size_t last = T->off.size() - 1;
  if (c.g >= end) {
    c.r = 1;
    if (c.k == last) return 0;  // the terminal 0: unique, smallest
    // separator 6, followed by the next contig's first base (< 6)
    return (uint64_t{6} << 41) | 1u;
  }
  const uint64_t s = code_at(T->words, c.g) + 1;
  const int64_t r = run_length(T, c.g, end);
  uint64_t b;
  if (c.g + r < end) b = code_at(T->words, c.g + r) + 1;
  else b = c.k == last ? 0 : 6;
  c.r = r;
  const bool longer_first = b > s;
  return (s << 41) | (uint64_t{longer_first} << 40) |
         (longer_first ? kMaxR - static_cast<uint64_t>(r) : static_cast<uint64_t>(r));
}

void advance(const bgo_sa_text* T, Cursor& c) {
  if (c.g >= T->off[c.k] + T->len[c.k]) {
    ++c.k;  // past the separator
    c.g = T->off[c.k];
  } else {
    c.g += c.r;
  }
}

template <class F>
void parallel_for(int64_t n, F f) {
  const int64_t kMin = int64_t{1} << 20;
  int nt = static_cast<int>(std::min<int64_t>(16, std::max<int64_t>(1, n / kMin)));
  nt = std::min<int>(nt, std::max(1u, std::thread::hardware_concurrency()));
  if (nt <= 1) {
    f(int64_t{0}, n);
    return;
  }
  std::vector<std::thread> ts;
  for (int t = 0; t < nt; ++t) ts.emplace_back(f, n * t / nt, n * (t + 1) / nt);
  for (auto& th : ts) th.join();
}

}  // namespace

extern "C" int bgo_sa_text_create(const uint32_t* words, int64_t n_words, int64_t total_bp,
                                  const int64_t* contig_offsets, const int64_t* contig_lengths,
                                  int32_t n_contigs, bgo_sa_text** out) {
  if (!out || !words || n_words <= 0 || total_bp <= 0 || total_bp > n_words * 16 ||
      total_bp >= (int64_t{1} << 32) || n_contigs <= 0 || !contig_offsets || !contig_lengths)
    return fail(BGO_ERR_ARG, "bgo_sa_text_create: bad arguments");
  int64_t expect = 0;
  for (int32_t k = 0; k < n_contigs; ++k) {
    if (contig_offsets[k] != expect || contig_lengths[k] <= 0)
      return fail(BGO_ERR_ARG, "bgo_sa_text_create: contigs must be non-empty and tile the genome");
    expect += contig_lengths[k];
  }
  if (expect != total_bp) return fail(BGO_ERR_ARG, "bgo_sa_text_create: contigs do not sum to total_bp");
  auto* T = new bgo_sa_text();
  T->words = words;
  T->total_bp = total_bp;
  T->off.assign(contig_offsets, contig_offsets + n_contigs);
  T->len.assign(contig_lengths, contig_lengths + n_contigs);
  find_long_runs(T);
  *out = T;
  return BGO_OK;
}

extern "C" int bgo_sa_text_destroy(bgo_sa_text* T) {
  delete T;
  return BGO_OK;
}

extern "C" int bgo_sa_text_long_runs(const bgo_sa_text* T, int64_t* n_out) {
  if (!T || !n_out) return fail(BGO_ERR_ARG, "null pointer");
  *n_out = static_cast<int64_t>(T->run_start.size());
  return BGO_OK;
}

extern "C" int bgo_sa_text_select(const bgo_sa_text* T, const uint32_t* positions, int64_t n,
                                  const uint8_t* anchor, int32_t m, int64_t cap, uint32_t* out,
                                  int64_t* n_occ_out, int64_t* n_out) {
  if (!T || n < 0 || (n > 0 && !positions) || !anchor || m <= 0 || m > 32 || cap < 1 ||
      !n_occ_out || !n_out || !out)
    return fail(BGO_ERR_ARG, "bgo_sa_text_select: bad arguments");
  // the anchor's occurrences among the positions: all m bases inside one
  // contig and equal to the anchor (the text's separators break a match)
  std::vector<uint8_t> ok(static_cast<size_t>(std::max<int64_t>(n, 0)), 0);
  parallel_for(n, [&](int64_t a, int64_t b) {
    for (int64_t i = a; i < b; ++i) {
      const int64_t q = positions[i];
      if (q >= T->total_bp) continue;
      const size_t k = static_cast<size_t>(
          std::upper_bound(T->off.begin(), T->off.end(), q) - T->off.begin()) - 1;
      if (q + m > T->off[k] + T->len[k]) continue;
      bool eq = true;
      for (int j = 0; j < m && eq; ++j) eq = code_at(T->words, q + j) == anchor[j];
      ok[static_cast<size_t>(i)] = eq;
    }
  });
  std::vector<uint32_t> occ;
  occ.reserve(static_cast<size_t>(std::max<int64_t>(n, 0)));
  for (int64_t i = 0; i < n; ++i)
    if (ok[static_cast<size_t>(i)]) occ.push_back(positions[i]);
  std::vector<uint8_t>().swap(ok);
  *n_occ_out = static_cast<int64_t>(occ.size());
  int64_t need = cap;
  int64_t got = 0;
  if (static_cast<int64_t>(occ.size()) <= need) {
    for (uint32_t q : occ) out[got++] = q;
    *n_out = got;
    return BGO_OK;
  }
  auto contig_of = [&](int64_t q) {
    return static_cast<uint32_t>(
        std::upper_bound(T->off.begin(), T->off.end(), q) - T->off.begin() - 1);
  };
  // level 0 over every occurrence: keys only, then a selection
  std::vector<uint64_t> keys(occ.size());
  parallel_for(static_cast<int64_t>(occ.size()), [&](int64_t a, int64_t b) {
    for (int64_t i = a; i < b; ++i) {
      Cursor c{occ[i], contig_of(occ[i]), occ[i], 0};
      keys[i] = run_key(T, c);
    }
  });
  std::vector<uint64_t> sel(keys);
  std::nth_element(sel.begin(), sel.begin() + (need - 1), sel.end());
  uint64_t kstar = sel[need - 1];
  std::vector<uint64_t>().swap(sel);
  std::vector<Cursor> group;
  for (size_t i = 0; i < occ.size(); ++i) {
    if (keys[i] < kstar) {
      out[got++] = occ[i];
    } else if (keys[i] == kstar) {
      Cursor c{occ[i], contig_of(occ[i]), occ[i], 0};
      run_key(T, c);
      advance(T, c);
      group.push_back(c);
    }
  }
  need -= got;
  std::vector<uint64_t>().swap(keys);
  std::vector<uint32_t>().swap(occ);
  // refine the group that straddles the cap; its cursors share a prefix
  while (need > 0 && !group.empty()) {
    if (static_cast<int64_t>(group.size()) <= need) {
      for (const Cursor& c : group) out[got++] = c.q;
      break;
    }
    std::vector<uint64_t> k2(group.size());
    for (size_t i = 0; i < group.size(); ++i) k2[i] = run_key(T, group[i]);
    std::vector<uint64_t> s2(k2);
    std::nth_element(s2.begin(), s2.begin() + (need - 1), s2.end());
    kstar = s2[need - 1];
    std::vector<Cursor> next;
    int64_t took = 0;
    for (size_t i = 0; i < group.size(); ++i) {
      if (k2[i] < kstar) {
        out[got++] = group[i].q;
        ++took;
      } else if (k2[i] == kstar) {
        Cursor c = group[i];
        advance(T, c);
        next.push_back(c);
      }
    }
    need -= took;
    group.swap(next);
  }
  std::sort(out, out + got);
  *n_out = got;
  return BGO_OK;
}
