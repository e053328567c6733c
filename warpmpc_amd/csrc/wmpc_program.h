// Host-side derivation of the kernels' row/column programs for one QDLDL
// pattern (gfx950 kernels, one lane per instance).
//
// The ragged QDLDLPlan tables (the padded tables of
// warpmpc/jax_qdldl/core.py:147-256 with the padding dropped) are turned into
// programs whose every operand is a wave-uniform table entry, and whose
// per-instance working values live in LDS "slots" ([slot][64 lanes] floats,
// lane-contiguous, so a wavefront's access to one slot is conflict-free):
//
//   * Y   — the dense scratch vector y of the up-looking factorization,
//           restricted to the reach of the current row (<= max reach slots);
//   * CW  — column window: Dinv[c] (and, when the forward substitution is fused
//           into the factorization, w[c]) from row c until the last row whose
//           reach contains c;
//   * AL  — live L entries: L[k,c] from row k until the last row that re-reads
//           column c, up to an LDS capacity; entries beyond it are re-read from
//           the caller's L array in global memory;
//   * W   — row window of the column-oriented sweeps: row r is live from the
//           first column that updates it (its smallest column, rowmin(r)) to
//           column r, in the forward (scatter) sweep and, in reverse, in the
//           backward (gather) sweep; both sweeps share one slot assignment.
//
// The sweeps walk L in CSC order, so their L addresses are arithmetic; the
// forward right-hand side is consumed in "first-touch" order (row r at its
// first update, or at column r if it has none), which lets the caller stage it
// as a sequential stream.
//
// Slots are assigned greedily in program order (interval colouring), so the
// slot counts equal the maximum number of simultaneously live values.
#pragma once

#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/warpmpc_hip.h"

namespace wmpc {

// flag bit of fent: this entry is the first update of its row
constexpr int32_t kFirstTouch = 1 << 16;

// Record streams of the chunked sweeps (wmpc_stream.h).  A record is two
// int32 words {w, a}; a chunk is kRecChunk records: two header records
// {l, v}, {d, nrec} giving the stream offsets of the chunk two ahead (L
// values, per-column values, D⁻¹ values) and this chunk's record count, then
// up to kRecChunk - 2 records:
//   entry  : w = W slot of li[q] (| kFirstTouch on its first update), a = 0
//   marker : w = kRecMark | W slot of row c (kRecNoSlot if none), a = p[c]
// (forward: the marker of column c precedes its entries; backward: entries
// in descending q, then the marker of their column).
constexpr int32_t kRecChunk = 64;
constexpr int32_t kRecMark = 1 << 20;
constexpr int32_t kRecNoSlot = 0xFFFF;
constexpr int32_t kRecSlotMask = 0xFFFF;
constexpr int32_t kRecVMax = 16;  // per-column values (rhs / D⁻¹ / stash) per chunk
constexpr int32_t kRecL = 64;     // L values per chunk (>= records per chunk)

struct HostProgram {
  int32_t n = 0, nnz_a = 0, nnz_l = 0;
  int32_t ymax = 0, ncol = 0, nact = 0, nbw = 0, act_need = 0;
  int32_t global_rereads = 0;  // L re-reads served from global memory
  // row k: {proc_end, init_end, colslot of k (-1 if column k is never reached), diag value slot}
  std::vector<int32_t> row;   // 4 per row
  std::vector<int32_t> init;  // 2 per entry: {yslot, original value slot}
  // reach entry r (row k, column c): {colslot(c) | (AL slot kept + 1) << 16 (0: not kept),
  // lidx of L[k,c], first update whose L comes from global memory, update end}
  std::vector<int32_t> proc;  // 4 per entry
  // update u: {target yslot, L source: AL slot >= 0, or -(lidx+1) for global}; the
  // updates of one reach entry hit distinct slots, so their order is free: those
  // reading a live-L slot come first, then the global re-reads
  std::vector<int32_t> upd;   // 2 per entry
  // column c (backward order): {bslot of c (-1 if row c has no entries), lp[c+1]}
  std::vector<int32_t> bcol;  // 2 per column
  std::vector<int32_t> bent;  // [nnz_l] bslot of li[q]
  std::vector<int32_t> p, pinv;
  // column sweeps: fcol[c] = {W slot of row c (-1: never updated), lp[c+1], 0, 0};
  // bcol2[i] (i = n-1-c) = {W slot of row c (-1 none), lp[c], p[c], 0};
  // fent[q] = W slot of li[q] | kFirstTouch if q is the first update of li[q];
  // ftpos[c] = position of row c in first-touch order.
  std::vector<int32_t> fcol, bcol2, fent, ftpos;
  std::vector<int32_t> rperm;  // rperm[ftpos[c]] = p[c]
  // chunked record streams (forward, backward) and the offsets of chunk 1
  std::vector<int32_t> frec, brec;
  int32_t fchunks = 0, bchunks = 0;
  int32_t f1[3] = {0, 0, 0}, b1[3] = {0, 0, 0};
};

// Packs a record list into chunks.  v_use[i] / d_use[i]: record i consumes a
// per-column value / a D⁻¹ value; l_use: it consumes an L value.
inline void pack_records(const std::vector<int32_t>& rec, const std::vector<uint8_t>& l_use,
                         const std::vector<uint8_t>& v_use, const std::vector<uint8_t>& d_use,
                         std::vector<int32_t>& out, int32_t& nchunks, int32_t first[3]) {
  const int32_t per = kRecChunk - 2;
  std::vector<int32_t> starts;  // record index where each chunk starts
  std::vector<int32_t> pl, pv, pd;
  int32_t l = 0, v = 0, d = 0;
  size_t i = 0;
  const size_t nr = rec.size() / 2;
  while (i < nr) {
    starts.push_back((int32_t)i);
    pl.push_back(l);
    pv.push_back(v);
    pd.push_back(d);
    int32_t cnt = 0, cv = 0, cd = 0;
    while (i < nr && cnt < per && cv + v_use[i] <= kRecVMax && cd + d_use[i] <= kRecVMax) {
      l += l_use[i];
      v += v_use[i];
      d += d_use[i];
      cv += v_use[i];
      cd += d_use[i];
      ++cnt;
      ++i;
    }
  }
  const int32_t real = (int32_t)starts.size();
  int32_t total = real + (real & 1) + 2;  // even, plus two empty chunks to prefetch
  starts.resize(total, (int32_t)nr);
  pl.resize(total, l);
  pv.resize(total, v);
  pd.resize(total, d);
  out.assign((size_t)total * kRecChunk * 2, 0);
  for (int32_t t = 0; t < total; ++t) {
    int32_t* c = &out[(size_t)t * kRecChunk * 2];
    const int32_t t2 = std::min(t + 2, total - 1);
    const int32_t b = starts[t], e = t + 1 < total ? starts[t + 1] : (int32_t)nr;
    c[0] = pl[t2];
    c[1] = pv[t2];
    c[2] = pd[t2];
    c[3] = e - b;
    for (int32_t k = b; k < e; ++k) {
      c[2 * (2 + k - b)] = rec[2 * (size_t)k];
      c[2 * (2 + k - b) + 1] = rec[2 * (size_t)k + 1];
    }
  }
  nchunks = total;
  first[0] = pl[1];
  first[1] = pv[1];
  first[2] = pd[1];
}

// Free-list slot allocator (lowest free slot first).
class SlotPool {
 public:
  explicit SlotPool(int32_t cap = INT32_MAX) : cap_(cap) {}
  int32_t take() {
    if (!free_.empty()) {
      auto it = std::min_element(free_.begin(), free_.end());
      int32_t s = *it;
      *it = free_.back();
      free_.pop_back();
      return s;
    }
    if (next_ >= cap_) return -1;
    return next_++;
  }
  void give(int32_t s) {
    if (s >= 0) free_.push_back(s);
  }
  int32_t high() const { return next_; }

 private:
  int32_t cap_;
  int32_t next_ = 0;
  std::vector<int32_t> free_;
};

// Builds the programs; returns an empty string on success, else the reason the
// pattern is inconsistent.  act_cap bounds the number of AL slots.
inline std::string build_program(const wmpc_qdldl_pattern* pt, int32_t act_cap, HostProgram& hp) {
  const int32_t n = pt->n;
  if (n <= 0 || pt->nnz_a <= 0 || pt->nnz_l < 0) return "pattern sizes must be positive";
  if (!(pt->perm_to_orig && pt->diag_pos && pt->init_ptr && pt->init_rows && pt->init_pos &&
        pt->proc_ptr && pt->proc_cols && pt->proc_lidx && pt->proc_off && pt->lp && pt->li && pt->p &&
        pt->pinv))
    return "pattern has a null table";
  if (pt->lp[0] != 0 || pt->lp[n] != pt->nnz_l) return "lp must span [0, nnz_l]";
  if (pt->init_ptr[0] != 0 || pt->proc_ptr[0] != 0) return "init_ptr/proc_ptr must start at 0";
  const int32_t n_proc = pt->proc_ptr[n];
  if (n_proc != pt->nnz_l) return "row programs must produce every entry of L once";
  for (int32_t i = 0; i < n; ++i) {
    if (pt->p[i] < 0 || pt->p[i] >= n || pt->pinv[pt->p[i]] != i) return "p/pinv mismatch";
    if (pt->lp[i + 1] < pt->lp[i]) return "lp must be nondecreasing";
  }
  for (int32_t q = 0; q < pt->nnz_l; ++q)
    if (pt->li[q] < 0 || pt->li[q] >= n) return "li out of range";
  hp = HostProgram();
  hp.n = n;
  hp.nnz_a = pt->nnz_a;
  hp.nnz_l = pt->nnz_l;
  hp.p.assign(pt->p, pt->p + n);
  hp.pinv.assign(pt->pinv, pt->pinv + n);

  // last(c): the last row whose reach contains c (== largest row index of L column c)
  std::vector<int32_t> last(n, -1);
  for (int32_t c = 0; c < n; ++c) {
    for (int32_t q = pt->lp[c]; q < pt->lp[c + 1]; ++q) {
      if (pt->li[q] <= c) return "L must be strictly lower triangular";
      if (q > pt->lp[c] && pt->li[q] <= pt->li[q - 1]) return "L row indices must increase";
    }
    if (pt->lp[c + 1] > pt->lp[c]) last[c] = pt->li[pt->lp[c + 1] - 1];
  }

  SlotPool colpool;
  SlotPool actpool(act_cap < 0 ? 0 : act_cap);
  std::vector<int32_t> colslot(n, -1), act_of(pt->nnz_l, -1), slot_of(n, -1), stamp(n, -1);
  std::vector<std::vector<int32_t>> ends_at(n);  // columns whose last row is k
  for (int32_t c = 0; c < n; ++c)
    if (last[c] >= 0) ends_at[last[c]].push_back(c);
  int32_t ymax = 0, act_live = 0, act_need = 0;
  std::vector<int32_t> act_need_delta(n + 1, 0);
  hp.row.resize(4 * (size_t)n);
  hp.init.resize(2 * (size_t)pt->init_ptr[n]);
  hp.proc.resize(4 * (size_t)n_proc);
  for (int32_t k = 0; k < n; ++k) {
    const int32_t dp = pt->diag_pos[k];
    if (dp < 0 || dp >= pt->nnz_a) return "diag_pos out of range";
    const int32_t r0 = pt->proc_ptr[k], r1 = pt->proc_ptr[k + 1];
    if (r1 < r0) return "proc_ptr must be nondecreasing";
    ymax = std::max(ymax, r1 - r0);
    for (int32_t r = r0; r < r1; ++r) {
      const int32_t c = pt->proc_cols[r];
      if (c < 0 || c >= k) return "reach column must precede its row";
      if (colslot[c] < 0) return "reach column has no window slot";
      slot_of[c] = r - r0;
      stamp[c] = k;
    }
    for (int32_t j = pt->init_ptr[k]; j < pt->init_ptr[k + 1]; ++j) {
      const int32_t row = pt->init_rows[j];
      if (row < 0 || row >= k || stamp[row] != k) return "init row outside the reach of its row";
      const int32_t ip = pt->init_pos[j];
      if (ip < 0 || ip >= pt->nnz_a) return "init_pos out of range";
      hp.init[2 * (size_t)j] = slot_of[row];
      hp.init[2 * (size_t)j + 1] = pt->perm_to_orig[ip];
    }
    for (int32_t r = r0; r < r1; ++r) {
      const int32_t c = pt->proc_cols[r];
      const int32_t off = pt->proc_off[r];
      if (off < 0 || pt->lp[c] + off >= pt->lp[c + 1]) return "proc_off out of range";
      if (pt->proc_lidx[r] != pt->lp[c] + off) return "proc_lidx must be lp[c] + offset";
      if (pt->li[pt->lp[c] + off] != k) return "L slot of a reach entry must hold row k";
      for (int32_t qq = 0; qq < off; ++qq) {
        const int32_t row = pt->li[pt->lp[c] + qq];
        if (row >= k || stamp[row] != k || slot_of[row] <= r - r0)
          return "update target must be a later reach slot";
      }
      for (int32_t qq = 0; qq < off; ++qq) {  // live-L operands
        const int32_t lidx = pt->lp[c] + qq;
        if (act_of[lidx] < 0) continue;
        hp.upd.push_back(slot_of[pt->li[lidx]]);
        hp.upd.push_back(act_of[lidx]);
      }
      const int32_t split = (int32_t)(hp.upd.size() / 2);
      for (int32_t qq = 0; qq < off; ++qq) {  // L entries re-read from global memory
        const int32_t lidx = pt->lp[c] + qq;
        if (act_of[lidx] >= 0) continue;
        hp.upd.push_back(slot_of[pt->li[lidx]]);
        hp.upd.push_back(-(lidx + 1));
        ++hp.global_rereads;
      }
      int32_t keep = -1;
      const int32_t lidx = pt->proc_lidx[r];
      if (last[c] > k) {  // re-read by a later row
        ++act_live;
        keep = actpool.take();
        act_of[lidx] = keep;
      }
      if (colslot[c] > 0xFFFF || keep >= 0xFFFF) return "window slot beyond the 16-bit record fields";
      int32_t* pe = &hp.proc[4 * (size_t)r];
      pe[0] = colslot[c] | ((keep + 1) << 16);
      pe[1] = lidx;
      pe[2] = split;
      pe[3] = (int32_t)(hp.upd.size() / 2);
    }
    act_need = std::max(act_need, act_live);
    // end of row k: release the windows of columns whose last row is k
    for (int32_t c : ends_at[k]) {
      colpool.give(colslot[c]);
      for (int32_t q = pt->lp[c]; q < pt->lp[c + 1]; ++q) {
        if (pt->li[q] >= k) break;
        actpool.give(act_of[q]);
        --act_live;
      }
    }
    if (last[k] > k) colslot[k] = colpool.take();
    int32_t* ri = &hp.row[4 * (size_t)k];
    ri[0] = r1;
    ri[1] = pt->init_ptr[k + 1];
    ri[2] = colslot[k];
    ri[3] = pt->perm_to_orig[dp];
  }
  hp.ymax = std::max(ymax, 1);
  hp.ncol = std::max(colpool.high(), 1);
  hp.nact = actpool.high();
  hp.act_need = act_need;

  // backward window, steps c = n-1 .. 0
  std::vector<int32_t> rowmin(n, INT32_MAX);
  for (int32_t c = 0; c < n; ++c)
    for (int32_t q = pt->lp[c]; q < pt->lp[c + 1]; ++q) rowmin[pt->li[q]] = std::min(rowmin[pt->li[q]], c);
  std::vector<std::vector<int32_t>> bend_at(n);
  for (int32_t r = 0; r < n; ++r)
    if (rowmin[r] != INT32_MAX) bend_at[rowmin[r]].push_back(r);
  SlotPool bpool;
  std::vector<int32_t> bslot(n, -1);
  hp.bcol.resize(2 * (size_t)n);
  hp.bent.resize(pt->nnz_l);
  for (int32_t c = n - 1; c >= 0; --c) {
    for (int32_t q = pt->lp[c]; q < pt->lp[c + 1]; ++q) {
      const int32_t r = pt->li[q];
      if (bslot[r] < 0) return "backward window slot missing";
      hp.bent[q] = bslot[r];
    }
    for (int32_t r : bend_at[c]) bpool.give(bslot[r]);
    if (rowmin[c] != INT32_MAX) bslot[c] = bpool.take();
    hp.bcol[2 * (size_t)c] = bslot[c];
    hp.bcol[2 * (size_t)c + 1] = pt->lp[c + 1];
  }
  hp.nbw = std::max(bpool.high(), 1);

  // column sweeps (forward scatter, backward gather) on the same slots
  hp.fcol.assign(4 * (size_t)n, 0);
  hp.bcol2.assign(4 * (size_t)n, 0);
  hp.fent.assign(pt->nnz_l, 0);
  hp.ftpos.assign(n, -1);
  int32_t pos = 0;
  for (int32_t c = 0; c < n; ++c) {
    hp.fcol[4 * (size_t)c] = bslot[c];
    hp.fcol[4 * (size_t)c + 1] = pt->lp[c + 1];
    if (rowmin[c] == INT32_MAX) hp.ftpos[c] = pos++;
    for (int32_t q = pt->lp[c]; q < pt->lp[c + 1]; ++q) {
      const int32_t r = pt->li[q];
      int32_t e = bslot[r];
      if (e < 0 || e >= kFirstTouch) return "row window slot out of range";
      if (rowmin[r] == c) {
        e |= kFirstTouch;
        hp.ftpos[r] = pos++;
      }
      hp.fent[q] = e;
    }
    const int32_t i = n - 1 - c;
    hp.bcol2[4 * (size_t)i] = bslot[c];
    hp.bcol2[4 * (size_t)i + 1] = pt->lp[c];
    hp.bcol2[4 * (size_t)i + 2] = pt->p[c];
  }
  if (pos != n) return "first-touch order does not cover every row";
  hp.rperm.assign(n, 0);
  for (int32_t c = 0; c < n; ++c) hp.rperm[hp.ftpos[c]] = pt->p[c];

  // record streams of the chunked sweeps
  {
    std::vector<int32_t> rec;
    std::vector<uint8_t> lu, vu, du;
    for (int32_t c = 0; c < n; ++c) {
      rec.push_back(kRecMark | (bslot[c] >= 0 ? bslot[c] : kRecNoSlot));
      rec.push_back(pt->p[c]);
      lu.push_back(0);
      vu.push_back(bslot[c] >= 0 ? 0 : 1);  // an unopened row takes its rhs here
      du.push_back(1);
      for (int32_t q = pt->lp[c]; q < pt->lp[c + 1]; ++q) {
        rec.push_back(hp.fent[q]);
        rec.push_back(0);
        lu.push_back(1);
        vu.push_back((hp.fent[q] & kFirstTouch) ? 1 : 0);
        du.push_back(0);
      }
    }
    pack_records(rec, lu, vu, du, hp.frec, hp.fchunks, hp.f1);
    rec.clear();
    lu.clear();
    vu.clear();
    du.clear();
    for (int32_t c = n - 1; c >= 0; --c) {
      for (int32_t q = pt->lp[c + 1] - 1; q >= pt->lp[c]; --q) {
        rec.push_back(hp.fent[q] & (kFirstTouch - 1));
        rec.push_back(0);
        lu.push_back(1);
        vu.push_back(0);
        du.push_back(0);
      }
      rec.push_back(kRecMark | (bslot[c] >= 0 ? bslot[c] : kRecNoSlot));
      rec.push_back(pt->p[c]);
      lu.push_back(0);
      vu.push_back(1);  // the D⁻¹-scaled forward value of column c
      du.push_back(0);
    }
    pack_records(rec, lu, vu, du, hp.brec, hp.bchunks, hp.b1);
  }
  return std::string();
}

}  // namespace wmpc
