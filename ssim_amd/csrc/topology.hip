// topology.hip -- what reset builds on the host and uploads once: the per-vertex and per-block contribution
// slots of the segmented reductions (global_composer.hpp:36-128 restated without maps), the symmetric
// blocked-ELL layout of the system matrix with its slot assignment, the per-tet material, the rest shapes and
// the Jacobi diagonal of the rest Hessian (finite_elements/time_step.hpp:368-494).
#include "context.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

using namespace ssim_hip;

namespace {

// Blocked-ELL slots of every block.  By default the upper blocks of a row take slots 0, 1, ... in column order
// (diagonal first) and the lower references likewise.  In the fixed 8/7 layout the slots are aligned by
// direction instead: slot k ≥ 1 holds, where it can, the block at the column offset w − v that sits at sorted
// position k in most full rows of its mesh (on a structured mesh every row then keeps each neighbour direction in
// the same slot), and the lower reference to the block stored in slot k of the neighbour's row goes to lower slot 7 − k.
// The fixed-width SpMV pairs upper slot k with lower slot 7 − k, so that for most rows of a slice the lower
// reads and the neighbouring slice's own reads of the same lines are issued at the same point of the program
// and meet in L2.  Slots only move blocks; every block is still read and summed once per row.
void assign_slots(ssim_ctx* c, const std::vector<int>& len) {
  const int nv = c->nv;
  c->bslot.assign(std::max(c->brow[nv], 1), 0);
  for (int v = 0; v < nv; ++v) {
    const int kd = c->kdiag[v];
    for (int k = 0; k < len[v]; ++k) c->bslot[c->brow[v] + k] = k < kd ? k : k - kd;
  }
  if (!c->fixed_layout) return;
  static_assert(kFixedWL == kFixedWU - 1, "the slot pairing needs one lower slot per off-diagonal upper slot");
  // Each mesh of a batch picks its own directions: the column offsets of a structured mesh's neighbours depend on
  // its cross-section, so one choice over a whole batch would align the slots of a few of its meshes only.
  std::vector<int> seg = c->mesh_rows;
  if (seg.size() < 2 || seg.front() != 0 || seg.back() != nv) seg = {0, nv};
  for (size_t g = 0; g + 1 < seg.size(); ++g) {
    const int r0 = seg[g], r1 = seg[g + 1];
    // modal column offset of each sorted upper position over the range's rows with all kFixedWU upper blocks
    std::vector<std::vector<int>> offs(kFixedWU);
    for (int v = r0; v < r1; ++v)
      if (len[v] - c->kdiag[v] == kFixedWU)
        for (int q = 1; q < kFixedWU; ++q) offs[q].push_back(c->bcol[c->brow[v] + c->kdiag[v] + q] - v);
    std::vector<long long> modal(kFixedWU, -1);  // −1: no canonical offset (offsets are positive)
    for (int q = 1; q < kFixedWU; ++q) {
      auto& o = offs[q];
      if (o.empty()) continue;
      std::sort(o.begin(), o.end());
      int best = o[0], run = 0, best_run = 0;
      for (size_t i = 0; i < o.size(); ++i) {
        run = (i > 0 && o[i] == o[i - 1]) ? run + 1 : 1;
        if (run > best_run) {
          best_run = run;
          best = o[i];
        }
      }
      modal[q] = best;
    }
    // upper slots: the diagonal in slot 0, canonical directions in their slot, the rest in the free slots
    for (int v = r0; v < r1; ++v) {
      const int kd = c->kdiag[v];
      bool used[kFixedWU] = {true};
      int rest[kFixedWU], nrest = 0;
      c->bslot[c->brow[v] + kd] = 0;
      for (int k = kd + 1; k < len[v]; ++k) {
        const long long off = c->bcol[c->brow[v] + k] - v;
        int q = 1;
        while (q < kFixedWU && !(modal[q] == off && !used[q])) ++q;
        if (q < kFixedWU) {
          c->bslot[c->brow[v] + k] = q;
          used[q] = true;
        } else {
          rest[nrest++] = k;
        }
      }
      for (int i = 0, q = 1; i < nrest; ++i) {
        while (used[q]) ++q;
        c->bslot[c->brow[v] + rest[i]] = q;
        used[q] = true;
      }
    }
  }
  // lower references: the block stored in slot k of row w goes to lower slot kFixedWU − 1 − k where free
  auto upper_slot_of = [&](int w, int v) -> int {
    const int* b = c->bcol.data() + c->brow[w];
    const int* e = c->bcol.data() + c->brow[w + 1];
    const int* it = std::lower_bound(b, e, v);
    if (it == e || *it != v) fail(SSIM_ERR_STATE, "internal: missing block in sparsity pattern");
    return c->bslot[c->brow[w] + (int)(it - b)];
  };
  for (int v = 0; v < nv; ++v) {
    const int kd = c->kdiag[v];
    bool used[kFixedWL] = {};
    int rest[kFixedWL], nrest = 0;
    for (int k = 0; k < kd; ++k) {
      const int j = kFixedWU - 1 - upper_slot_of(c->bcol[c->brow[v] + k], v);
      if (j >= 0 && j < kFixedWL && !used[j]) {
        c->bslot[c->brow[v] + k] = j;
        used[j] = true;
      } else {
        rest[nrest++] = k;
      }
    }
    for (int i = 0, j = 0; i < nrest; ++i) {
      while (used[j]) ++j;
      c->bslot[c->brow[v] + rest[i]] = j;
      used[j] = true;
    }
  }
}

// Sparsity, blocked-ELL layout and the two assembly permutations.
void build_topology(ssim_ctx* c) {
  const int nv = c->nv, nt = c->nt;
  const int* cells = c->hcells.data();
  // block adjacency: neighbours (self included, also for a vertex in no tet) of every vertex, ascending.
  // Tets with a negative first index are absent (padding of a batched layout) and take no part.
  std::vector<int> cnt(nv + 1, 1);
  cnt[0] = 0;
  for (int e = 0; e < nt; ++e)
    if (cells[4 * e] >= 0)
      for (int i = 0; i < 4; ++i) cnt[cells[4 * e + i] + 1] += 4;
  for (int v = 0; v < nv; ++v) cnt[v + 1] += cnt[v];
  std::vector<int> cand(std::max(cnt[nv], 1));
  std::vector<int> fillp(cnt.begin(), cnt.end() - 1);
  for (int v = 0; v < nv; ++v) cand[fillp[v]++] = v;
  for (int e = 0; e < nt; ++e)
    if (cells[4 * e] >= 0)
      for (int i = 0; i < 4; ++i) {
        const int v = cells[4 * e + i];
        for (int j = 0; j < 4; ++j) cand[fillp[v]++] = cells[4 * e + j];
      }
  c->brow.assign(nv + 1, 0);
  std::vector<int> len(nv, 0);
  for (int v = 0; v < nv; ++v) {
    int* b = cand.data() + cnt[v];
    int* e = cand.data() + cnt[v + 1];
    std::sort(b, e);
    len[v] = (int)(std::unique(b, e) - b);
    c->brow[v + 1] = c->brow[v] + len[v];
  }
  c->bcol.resize(std::max(c->brow[nv], 1));
  for (int v = 0; v < nv; ++v) std::copy(cand.begin() + cnt[v], cand.begin() + cnt[v] + len[v], c->bcol.begin() + c->brow[v]);
  cand.clear();
  cand.shrink_to_fit();
  auto slot = [&](int v, int w) -> int {  // index of w in the sorted neighbour list of v
    const int* b = c->bcol.data() + c->brow[v];
    const int* e = c->bcol.data() + c->brow[v + 1];
    const int* it = std::lower_bound(b, e, w);
    if (it == e || *it != w) fail(SSIM_ERR_STATE, "internal: missing block in sparsity pattern");
    return (int)(it - b);
  };
  c->kdiag.resize(nv);
  for (int v = 0; v < nv; ++v) c->kdiag[v] = slot(v, v);

  // symmetric blocked ELL: slices of 64 rows; the upper part (w ≥ v, diagonal first) is stored, the lower
  // part refers to the stored transposes
  c->nslices = (nv + kSlice - 1) / kSlice;
  const int ns = std::max(c->nslices, 1);
  c->hwu.assign(ns, 0);
  c->hoffu.assign(ns + 1, 0);
  std::vector<int> hwl(ns, 0), hoffl(ns + 1, 0);
  int max_wu = 0, max_wl = 0;
  for (int s = 0; s < c->nslices; ++s) {
    for (int l = 0; l < kSlice && s * kSlice + l < nv; ++l) {
      const int v = s * kSlice + l;
      c->hwu[s] = std::max(c->hwu[s], len[v] - c->kdiag[v]);
      hwl[s] = std::max(hwl[s], c->kdiag[v]);
    }
    max_wu = std::max(max_wu, c->hwu[s]);
    max_wl = std::max(max_wl, hwl[s]);
  }
  // rows narrow enough for the fixed layout, with every column within kBellMaxOffset vertices of its row: pad
  // every slice to the same widths
  int max_off = 0;
  for (int v = 0; v < nv; ++v)
    if (len[v] > 0)
      max_off = std::max({max_off, c->bcol[c->brow[v + 1] - 1] - v, v - c->bcol[c->brow[v]]});
  c->fixed_layout = max_wu <= kFixedWU && max_wl <= kFixedWL && max_off <= kBellMaxOffset &&
                    (long long)kBellZeroLead + (long long)ns * kBellSliceVals < (1LL << 31);
  if (c->fixed_layout)
    for (int s = 0; s < c->nslices; ++s) {
      c->hwu[s] = kFixedWU;
      hwl[s] = kFixedWL;
    }
  // each layout has its own row products: the fixed layout the unrolled one (0), the general one 1 or 4 lanes
  if (c->fixed_layout)
    c->spmv_lanes = 0;
  else if (c->spmv_lanes == 0 || c->spmv_lanes_default)
    c->spmv_lanes = 1;
  assign_slots(c, len);
  long long offu = 0, offl = 0;
  for (int s = 0; s < c->nslices; ++s) {
    c->hoffu[s] = (int)offu;
    hoffl[s] = (int)offl;
    offu += (long long)c->hwu[s] * kSlice;
    offl += (long long)hwl[s] * kSlice;
  }
  c->ncolslots = offu;
  c->nlowslots = offl;

  if (9 * offu + 9 * kSlice >= (1LL << 31) || offl >= (1LL << 31))
    fail(SSIM_ERR_INVALID_ARGUMENT, "mesh too large for 32-bit blocked-ELL offsets");
  auto valpos = [&](int v, int ku) -> int { return (int)c->valpos(v, ku); };
  std::vector<int> hucol;
  std::vector<int2> hlref;
  std::vector<uint4> hidx;
  if (c->fixed_layout) {
    // two index words per row: 16-bit column offsets of upper slots 1..7, 16-bit row offsets of lower slots
    // 0..6 and their partners' upper slots (3 bits each); padding stays 0 (column v with zero values; the
    // leading zero block)
    std::vector<unsigned> up(7 * (size_t)ns * kSlice, 0), lo(7 * (size_t)ns * kSlice, 0), sw((size_t)ns * kSlice, 0);
    for (int v = 0; v < nv; ++v) {
      const int kd = c->kdiag[v];
      for (int k = kd + 1; k < len[v]; ++k) {
        const int ku = c->bslot[c->brow[v] + k];
        up[7 * (size_t)v + ku - 1] = (unsigned)(c->bcol[c->brow[v] + k] - v);
      }
      for (int k = 0; k < kd; ++k) {
        const int w = c->bcol[c->brow[v] + k];
        const int j = c->bslot[c->brow[v] + k];
        const int kw = c->bslot[c->brow[w] + slot(w, v)];
        lo[7 * (size_t)v + j] = (unsigned)(v - w);
        sw[v] |= (unsigned)kw << (3 * j);
      }
    }
    hidx.assign(2 * (size_t)ns * kSlice, make_uint4(0, 0, 0, 0));
    for (int v = 0; v < nv; ++v) {
      const unsigned* u = &up[7 * (size_t)v];
      const unsigned* l = &lo[7 * (size_t)v];
      const int s = v / kSlice, lane = v % kSlice;
      hidx[(2 * (size_t)s) * kSlice + lane] =
          make_uint4(u[0] | u[1] << 16, u[2] | u[3] << 16, u[4] | u[5] << 16, u[6] | l[0] << 16);
      hidx[(2 * (size_t)s + 1) * kSlice + lane] =
          make_uint4(l[1] | l[2] << 16, l[3] | l[4] << 16, l[5] | l[6] << 16, sw[v]);
    }
  } else {
    // padding: upper slots → (v, zero values); lower references → (v, a zero block after the last slice)
    hucol.assign(std::max<long long>(offu, 1), 0);
    hlref.assign(std::max<long long>(offl, 1), make_int2(0, 0));
    const long long zero_block = 9 * offu;
    for (int s = 0; s < c->nslices; ++s)
      for (int l = 0; l < kSlice; ++l) {
        const int v = std::min(s * kSlice + l, nv - 1);
        for (int k = 0; k < c->hwu[s]; ++k) hucol[c->hoffu[s] + k * kSlice + l] = v;
        for (int k = 0; k < hwl[s]; ++k) hlref[hoffl[s] + k * kSlice + l] = make_int2(v, (int)(zero_block + l));
      }
    for (int v = 0; v < nv; ++v) {
      const int s = v / kSlice, lane = v % kSlice;
      const int kd = c->kdiag[v];
      for (int k = kd; k < len[v]; ++k)
        hucol[c->hoffu[s] + c->bslot[c->brow[v] + k] * kSlice + lane] = c->bcol[c->brow[v] + k];
      for (int k = 0; k < kd; ++k) {
        const int w = c->bcol[c->brow[v] + k];
        hlref[hoffl[s] + c->bslot[c->brow[v] + k] * kSlice + lane] =
            make_int2(w, valpos(w, c->bslot[c->brow[w] + slot(w, v)]));
      }
    }
  }

  // upper blocks (r ≤ c) in row-major order, with the value base of (r,c)
  std::vector<int> ubptr(nv + 1, 0);
  for (int v = 0; v < nv; ++v) ubptr[v + 1] = ubptr[v] + (len[v] - c->kdiag[v]);
  c->nub = ubptr[nv];
  std::vector<int2> hub_rc(std::max(c->nub, 1));
  std::vector<int> hub_pos(std::max(c->nub, 1));
  for (int v = 0; v < nv; ++v)
    for (int k = c->kdiag[v]; k < len[v]; ++k) {
      const int ub = ubptr[v] + (k - c->kdiag[v]);
      hub_rc[ub] = make_int2(v, c->bcol[c->brow[v] + k]);
      const int pos = valpos(v, c->bslot[c->brow[v] + k]);
      hub_pos[ub] = (c->fixed_layout && k == c->kdiag[v]) ? -(pos + 1) : pos;  // 6-value diagonal
    }

  // Hessian contributions: 10 local pairs (v ≤ w) per tet, sorted by (upper block, tet, pair)
  std::vector<int> hcount(c->nub + 1, 0);
  std::vector<int> hub_of(10 * (size_t)nt, -1);
  for (int e = 0; e < nt; ++e)
    if (cells[4 * e] >= 0)
      for (int w = 0; w < 4; ++w)
        for (int v = 0; v <= w; ++v) {
          const int a = cells[4 * e + v], b = cells[4 * e + w];
          const int r = std::min(a, b), cc = std::max(a, b);
          const int ub = ubptr[r] + (slot(r, cc) - c->kdiag[r]);
          hub_of[10 * (size_t)e + w * (w + 1) / 2 + v] = ub;
          hcount[ub + 1]++;
        }
  for (int u = 0; u < c->nub; ++u) hcount[u + 1] += hcount[u];
  // the same blocks with each tet's corners in ascending vertex order (the tile-reduced assembly's pair order)
  {
    std::vector<int> hub_sorted(10 * (size_t)nt, -1);
    for (int e = 0; e < nt; ++e) {
      if (cells[4 * e] < 0) continue;
      int id[4] = {cells[4 * e], cells[4 * e + 1], cells[4 * e + 2], cells[4 * e + 3]};
      std::sort(id, id + 4);
      for (int w = 0; w < 4; ++w)
        for (int v = 0; v <= w; ++v)
          hub_sorted[10 * (size_t)e + w * (w + 1) / 2 + v] = ubptr[id[v]] + (slot(id[v], id[w]) - c->kdiag[id[v]]);
    }
    build_hess_tiles(c, hub_sorted);
  }
  std::vector<int> hperm(10 * (size_t)nt, -1);
  {
    std::vector<int> pos(hcount.begin(), hcount.end() - 1);
    for (size_t k = 0; k < 10 * (size_t)nt; ++k)
      if (hub_of[k] >= 0) hperm[k] = pos[hub_of[k]]++;
  }
  // gradient contributions: 4 per tet, sorted by (vertex, tet, corner)
  std::vector<int> gseg(nv + 1, 0);
  for (int e = 0; e < nt; ++e)
    if (cells[4 * e] >= 0)
      for (int i = 0; i < 4; ++i) gseg[cells[4 * e + i] + 1]++;
  for (int v = 0; v < nv; ++v) gseg[v + 1] += gseg[v];
  std::vector<int> gperm(4 * (size_t)nt, -1);
  {
    std::vector<int> pos(gseg.begin(), gseg.end() - 1);
    for (int e = 0; e < nt; ++e)
      if (cells[4 * e] >= 0)
        for (int i = 0; i < 4; ++i) gperm[4 * e + i] = pos[cells[4 * e + i]]++;
  }
  hipStream_t s = c->stream;
  build_grad_tiles(c);
  c->wu.upload(c->hwu.data(), ns, s);
  c->offu.upload(c->hoffu.data(), ns, s);
  c->wl.upload(hwl.data(), ns, s);
  c->offl.upload(hoffl.data(), ns, s);
  if (c->fixed_layout) {
    c->ucol.release();
    c->lref.release();
    c->idx.upload(hidx, s);
  } else {
    c->idx.release();
    c->ucol.upload(hucol, s);
    c->lref.upload(hlref, s);
  }
  c->ub_rc.upload(hub_rc, s);
  c->ub_pos.upload(hub_pos, s);
  c->hseg.upload(hcount, s);
  c->hperm.upload(hperm, s);
  c->gseg.upload(gseg, s);
  c->gperm.upload(gperm, s);
  // + the zero block of the padded lower references (after the last slice, or leading in the fixed layout)
  c->val.alloc(c->fixed_layout ? kBellZeroLead + (size_t)ns * kBellSliceVals : 9 * (size_t)offu + 9 * kSlice);
  HIPCHK(hipMemsetAsync(c->val.p, 0, c->val.count * sizeof(double), s));
  c->gcontrib.alloc(12 * (size_t)std::max(nt, 1));
  c->hcontrib.release();  // per-tet contribution slots: allocated on first use (projective dynamics, hess_tile = 0)
  sync(c);  // host staging vectors go out of scope
}

}  // namespace

// Tiles of kHessTile consecutive tets for the tile-reduced Hessian: per tile its upper blocks ascending, the tile's
// (tet, pair) contributions sorted by (upper block, tet, pair) into LDS slots, and the run boundaries of every tile
// block; per upper block the positions of its tile partials in tile order.  hub[10e + k] is the upper block of pair
// k = w(w+1)/2 + v of tet e with its corners in ascending vertex order (−1 for a padding tet).
void ssim_hip::build_hess_tiles(ssim_ctx* c, const std::vector<int>& hub) {
  const int nt = c->nt, T = kHessTile;
  const int ntiles = (nt + T - 1) / T;
  std::vector<int> ht_off(ntiles + 1, 0), ht_run, ht_blk;
  std::vector<unsigned> ht_slot(5 * (size_t)std::max(nt, 1), 0u);
  std::vector<std::pair<int, int>> keyed;  // (upper block, contribution index 10·(e − e0) + k)
  std::vector<int> slot_of(10 * T);
  for (int t = 0; t < ntiles; ++t) {
    const int e0 = t * T, e1 = std::min(nt, e0 + T);
    keyed.clear();
    for (int e = e0; e < e1; ++e)
      for (int k = 0; k < 10; ++k)
        if (hub[10 * (size_t)e + k] >= 0) keyed.emplace_back(hub[10 * (size_t)e + k], 10 * (e - e0) + k);
    std::sort(keyed.begin(), keyed.end());
    int nb = 0;
    for (size_t q = 0; q < keyed.size(); ++q) {
      if (q == 0 || keyed[q].first != keyed[q - 1].first) {
        ht_blk.push_back(keyed[q].first);
        ht_run.push_back((int)q);
        ++nb;
      }
      slot_of[keyed[q].second] = (int)q;
    }
    ht_run.push_back((int)keyed.size());  // nb + 1 boundaries
    ht_off[t + 1] = ht_off[t] + nb;
    for (int e = e0; e < e1; ++e)
      for (int k = 0; k < 10; ++k)
        if (hub[10 * (size_t)e + k] >= 0)
          ht_slot[(size_t)(k >> 1) * nt + e] |= (unsigned)slot_of[10 * (e - e0) + k] << (16 * (k & 1));
  }
  std::vector<int> hb_seg(c->nub + 1, 0), hb_pos(std::max<size_t>(ht_blk.size(), 1), 0);
  for (int b : ht_blk) hb_seg[b + 1]++;
  for (int u = 0; u < c->nub; ++u) hb_seg[u + 1] += hb_seg[u];
  {
    std::vector<int> pos(hb_seg.begin(), hb_seg.end() - 1);
    for (size_t k = 0; k < ht_blk.size(); ++k) hb_pos[pos[ht_blk[k]]++] = (int)k;  // k ascends with the tile
  }
  hipStream_t s = c->stream;
  c->ht_slot.upload(ht_slot, s);
  c->ht_off.upload(ht_off, s);
  if (ht_run.empty()) ht_run.push_back(0);
  c->ht_run.upload(ht_run, s);
  c->hb_seg.upload(hb_seg, s);
  c->hb_pos.upload(hb_pos, s);
  c->hparts = (long long)ht_blk.size();
  c->hpart.alloc(9 * (size_t)std::max<long long>(c->hparts, 1));
  sync(c);  // host staging vectors go out of scope
}

// Tiles of grad_tile consecutive tets for the tile-reduced gradient: per tile its vertices ascending, the tile's
// corners sorted by (vertex, tet, corner) into LDS slots, and the run boundaries of every tile vertex; per vertex
// the positions of its tile partials in tile order.
void ssim_hip::build_grad_tiles(ssim_ctx* c) {
  const int nv = c->nv, nt = c->nt, T = c->grad_tile;
  const int* cells = c->hcells.data();
  const int ntiles = (nt + T - 1) / T;
  std::vector<int> gt_off(ntiles + 1, 0), gt_run, gt_vert;
  std::vector<ushort4> gt_slot(std::max(nt, 1), make_ushort4(0, 0, 0, 0));
  {
    std::vector<int> tv, cnt, fill;
    for (int t = 0; t < ntiles; ++t) {
      const int e0 = t * T, e1 = std::min(nt, e0 + T);
      tv.clear();
      for (int e = e0; e < e1; ++e)
        if (cells[4 * e] >= 0)
          for (int i = 0; i < 4; ++i) tv.push_back(cells[4 * e + i]);
      std::sort(tv.begin(), tv.end());
      tv.erase(std::unique(tv.begin(), tv.end()), tv.end());
      const int ntv = (int)tv.size();
      cnt.assign(ntv + 1, 0);
      for (int e = e0; e < e1; ++e)
        if (cells[4 * e] >= 0)
          for (int i = 0; i < 4; ++i) cnt[std::lower_bound(tv.begin(), tv.end(), cells[4 * e + i]) - tv.begin() + 1]++;
      for (int j = 0; j < ntv; ++j) cnt[j + 1] += cnt[j];
      gt_run.insert(gt_run.end(), cnt.begin(), cnt.end());  // ntv + 1 boundaries
      fill.assign(cnt.begin(), cnt.end() - 1);
      for (int e = e0; e < e1; ++e) {
        if (cells[4 * e] < 0) continue;
        unsigned short sl[4];
        for (int i = 0; i < 4; ++i)
          sl[i] = (unsigned short)fill[std::lower_bound(tv.begin(), tv.end(), cells[4 * e + i]) - tv.begin()]++;
        gt_slot[e] = make_ushort4(sl[0], sl[1], sl[2], sl[3]);
      }
      gt_off[t + 1] = gt_off[t] + ntv;
      gt_vert.insert(gt_vert.end(), tv.begin(), tv.end());
    }
  }
  std::vector<int> gv_seg(nv + 1, 0), gv_pos(std::max<size_t>(gt_vert.size(), 1), 0);
  for (int v : gt_vert) gv_seg[v + 1]++;
  for (int v = 0; v < nv; ++v) gv_seg[v + 1] += gv_seg[v];
  {
    std::vector<int> pos(gv_seg.begin(), gv_seg.end() - 1);
    for (size_t k = 0; k < gt_vert.size(); ++k) gv_pos[pos[gt_vert[k]]++] = (int)k;  // k ascends with the tile
  }
  hipStream_t s = c->stream;
  c->gt_slot.upload(gt_slot, s);
  c->gt_off.upload(gt_off, s);
  if (gt_run.empty()) gt_run.push_back(0);
  c->gt_run.upload(gt_run, s);
  c->gv_seg.upload(gv_seg, s);
  c->gv_pos.upload(gv_pos, s);
  c->gpart.alloc(3 * std::max<size_t>(gt_vert.size(), 1));
  sync(c);  // host staging vectors go out of scope
}

void ssim_hip::upload_material(ssim_ctx* c) {
  std::vector<double> lam(c->nt), mu(c->nt);
  if (c->per_tet) {
    lam = c->hlam;
    mu = c->hmu;
  } else {
    // compute_lambda / compute_mu: elast/basic_elast.hpp:11-20
    const double l = c->youngs * c->poisson / ((1 + c->poisson) * (1 - 2 * c->poisson));
    const double m = c->youngs / (2 * (1 + c->poisson));
    std::fill(lam.begin(), lam.end(), l);
    std::fill(mu.begin(), mu.end(), m);
  }
  // a material that is the same (λ, μ) bitwise in every tet is passed to the element kernels as two scalars, so
  // that they read no per-tet material arrays
  c->uniform_material = true;
  for (int e = 1; e < c->nt && c->uniform_material; ++e)
    c->uniform_material = lam[e] == lam[0] && mu[e] == mu[0];
  c->lam0 = c->nt ? lam[0] : 0.0;
  c->mu0 = c->nt ? mu[0] : 0.0;
  c->lam.upload(lam, c->stream);
  c->mu.upload(mu, c->stream);
  sync(c);
}

// reset(): finite_elements/time_step.hpp:368-494

void ssim_hip::do_reset(ssim_ctx* c) {
  check_material(c->youngs, c->poisson, c->density);
  if (0.5 - c->poisson < 1e-4)
    fprintf(stderr, "Warning: got extremely high Poisson's ratio, may cause numerical instability. %lf\n", c->poisson);
  hipStream_t s = c->stream;
  const int n = c->n;
  c->rho_reset = c->density;
  // vertex buffers: inertia, next, forces zeroed; the others kept if already allocated (:384-413)
  for (int w = 0; w < SSIM_ARRAY_COUNT; ++w) {
    const bool keep = (w == SSIM_ARRAY_EXT_ACCEL || w == SSIM_ARRAY_DBC_VALUES || w == SSIM_ARRAY_PREV_DEFORM ||
                       w == SSIM_ARRAY_DEFORMATION || w == SSIM_ARRAY_VELOCITY) &&
                      c->arr[w].p && c->arr[w].count == (size_t)n;
    if (!keep) {
      c->arr[w].alloc(n);
      HIPCHK(hipMemsetAsync(c->arr[w].p, 0, sizeof(double) * std::max(n, 1), s));
    }
  }
  if (!(c->dof.p && c->dof.count == (size_t)n)) {
    c->dof.alloc(n);
    HIPCHK(hipMemsetAsync(c->dof.p, 0, std::max(n, 1), s));
  }
  for (DBuf<double>* b : {&c->utry, &c->gtry, &c->dir, &c->r, &c->z, &c->p, &c->q, &c->tmp, &c->dinv}) {
    b->alloc(n);
    HIPCHK(hipMemsetAsync(b->p, 0, sizeof(double) * std::max(n, 1), s));
  }
  c->hrest.release();
  c->B.alloc(9 * (size_t)std::max(c->nt, 1));
  c->V.alloc(std::max(c->nt, 1));
  upload_material(c);
  launch_rest_shape(c->X.p, c->cells.p, c->nt, c->B.p, c->V.p, s);
  HIPCHK(hipGetLastError());
  build_topology(c);
  // the SpMV grid has one workgroup per 64-row slice (or per 16 rows with 4 lanes per row), rounded up to a
  // multiple of 8 (XCD remapping)
  c->pmax = std::max<size_t>((size_t)max_partials(std::max({c->nt, c->n, c->nub, 1})), 4 * (size_t)c->nslices) + 8;
  c->partials.alloc(8 * c->pmax);
  // the projective-dynamics matrix is built at reset (projective_dynamics.hpp:51-87), for that stepper
  c->pd_val.release();
  if (c->stepper == SSIM_STEPPER_LBFGS_PD) build_pd(c);
  // the preconditioner factor is rebuilt on the new pattern and set to the identity (fem_cuda.cu:175-201)
  for (ssim_hip::DBuf<int>* b : {&c->pre_brow, &c->pre_bcol, &c->pre_tpos}) b->release();
  if (c->pre_val.p || c->stepper == SSIM_STEPPER_NCG_EXT_PREC) set_precond(c, nullptr);
  c->drop_graphs();  // buffers may have moved
  c->pcg.alloc(1);
  c->pcg_it.alloc(1);
  c->ev.alloc(1);
  create_poll_events(c);
  if (c->prm.pcg_precond == 1) ensure_unit_diagonal(c);
  HIPCHK(hipMemsetAsync(c->pcg.p, 0, sizeof(PcgScalars), s));
  HIPCHK(hipMemsetAsync(c->pcg_it.p, 0, sizeof(PcgIter), s));
  HIPCHK(hipMemsetAsync(c->ev.p, 0, sizeof(EvalScalars), s));
  // total mass Σ ρV (= asum(M.values)/3, time_step.hpp:490)
  std::vector<double> hV(c->nt);
  if (c->nt) HIPCHK(hipMemcpyAsync(hV.data(), c->V.p, sizeof(double) * c->nt, hipMemcpyDeviceToHost, s));
  sync(c);
  double mass = 0;
  for (int e = 0; e < c->nt; ++e) mass += c->density * hV[e];
  c->total_mass = mass;
  c->tol_abs = c->rtol * mass;
  c->has_reset = true;
}

