// Host-side construction of the EBE block layout (see layout.h).
#include "layout.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <numeric>

namespace bhip {

namespace {

// Hilbert index of (x, y) on a 2^order grid (classic rotate-and-flip walk).
uint64_t hilbert_d(uint32_t n, uint32_t x, uint32_t y) {
  uint64_t d = 0;
  for (uint32_t s = n / 2; s > 0; s /= 2) {
    uint32_t rx = (x & s) > 0;
    uint32_t ry = (y & s) > 0;
    d += (uint64_t)s * s * ((3 * rx) ^ ry);
    if (ry == 0) {
      if (rx == 1) {
        x = s - 1 - x;
        y = s - 1 - y;
      }
      std::swap(x, y);
    }
  }
  return d;
}

std::vector<int32_t> order_nodes(int npoin2, const double* x, const double* y) {
  std::vector<int32_t> order(npoin2);
  std::iota(order.begin(), order.end(), 0);
  if (!x || !y) return order;
  double xmin = x[0], xmax = x[0], ymin = y[0], ymax = y[0];
  for (int i = 1; i < npoin2; ++i) {
    xmin = std::min(xmin, x[i]);
    xmax = std::max(xmax, x[i]);
    ymin = std::min(ymin, y[i]);
    ymax = std::max(ymax, y[i]);
  }
  const uint32_t n = 1u << 16;
  double span = std::max(xmax - xmin, ymax - ymin);
  if (!(span > 0)) return order;
  std::vector<uint64_t> key(npoin2);
  for (int i = 0; i < npoin2; ++i) {
    uint32_t ix = (uint32_t)std::min<double>(n - 1, (x[i] - xmin) / span * (n - 1));
    uint32_t iy = (uint32_t)std::min<double>(n - 1, (y[i] - ymin) / span * (n - 1));
    key[i] = hilbert_d(n, ix, iy);
  }
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return key[a] < key[b]; });
  return order;
}

}  // namespace

std::string build_layout(int npoin, int nelem, int ndp, const int32_t* ikle,
                         const double* x, const double* y, int nplan,
                         int block_nodes, Layout& L) {
  if (npoin <= 0 || nelem <= 0) return "empty mesh";
  if (ndp != 2 && ndp != 3 && ndp != 6) return "ndp must be 2, 3 or 6";
  if (nplan < 1) nplan = 1;
  if (nplan > 1 && npoin % nplan != 0) return "npoin is not a multiple of nplan";
  for (long i = 0; i < (long)ndp * nelem; ++i)
    if (ikle[i] < 0 || ikle[i] >= npoin) return "connectivity out of range";
  for (int e = 0; e < nelem; ++e)
    for (int a = 0; a < ndp; ++a)
      for (int b = a + 1; b < ndp; ++b)
        if (ikle[(long)a * nelem + e] == ikle[(long)b * nelem + e])
          return "element with a repeated node";

  L = Layout();
  L.npoin = npoin;
  L.nelem = nelem;
  L.ndp = ndp;
  L.nplan = nplan;

  // ---- node order: Hilbert on the 2D base, prisms column by column ------
  const int npoin2 = npoin / nplan;
  std::vector<int32_t> ord2 = order_nodes(npoin2, x, y);
  L.perm.resize(npoin);
  L.iperm.resize(npoin);
  for (int r = 0; r < npoin2; ++r)
    for (int p = 0; p < nplan; ++p) {
      int internal = r * nplan + p;
      int caller = p * npoin2 + ord2[r];
      L.perm[internal] = caller;
    }
  for (int i = 0; i < npoin; ++i) L.iperm[L.perm[i]] = i;

  // sigma-mesh prisms (top node k + 3 = bottom node k one plane up, caller
  // node p * npoin2 + i): columns are numbered contiguously, so the EBE
  // product reads the three bottom ids of a prism and derives the top ones
  if (ndp == 6 && nplan > 1) {
    bool cols = true;
    for (int e = 0; e < nelem && cols; ++e)
      for (int k = 0; k < 3; ++k)
        if (ikle[(long)(k + 3) * nelem + e] != ikle[(long)k * nelem + e] + npoin2) cols = false;
    L.col_ids = cols;
  }

  int nb = block_nodes > 0 ? block_nodes : 1024;
  if (nplan > 1) nb = std::max(nplan, (nb / nplan) * nplan);
  L.block_nodes = nb;
  L.nblocks = (npoin + nb - 1) / nb;
  L.node_start.resize(L.nblocks + 1);
  for (int b = 0; b <= L.nblocks; ++b) L.node_start[b] = std::min(npoin, b * nb);

  // ---- node -> element adjacency (internal node ids) --------------------
  std::vector<int32_t> deg(npoin + 1, 0);
  for (long i = 0; i < (long)ndp * nelem; ++i) deg[L.iperm[ikle[i]] + 1]++;
  std::vector<int64_t> adj_off(npoin + 1, 0);
  for (int i = 0; i < npoin; ++i) adj_off[i + 1] = adj_off[i] + deg[i + 1];
  std::vector<int32_t> adj(adj_off[npoin]);
  {
    std::vector<int64_t> fill(adj_off.begin(), adj_off.end() - 1);
    for (int e = 0; e < nelem; ++e)
      for (int k = 0; k < ndp; ++k) adj[fill[L.iperm[ikle[(long)k * nelem + e]]]++] = e;
  }

  // ---- per block: element instances, halo, colours ----------------------
  std::vector<int32_t> elem_stamp(nelem, -1);
  std::vector<int32_t> node_stamp(npoin, -1);
  std::vector<int32_t> node_local(npoin, 0);
  std::vector<uint64_t> owned_mask;  // colour bitmask per owned node
  std::vector<int32_t> belems, bcolor, sorted;
  std::vector<std::vector<uint16_t>> lid_cols(ndp);
  L.halo_start.push_back(0);
  L.color_off.assign((size_t)L.nblocks * (kMaxColors + 1), 0);
  L.ncolor.assign(L.nblocks, 0);
  std::vector<int32_t> slot_elem;
  slot_elem.reserve((size_t)(nelem * 1.3) + 16);
  for (auto& c : lid_cols) c.reserve((size_t)(nelem * 1.3) + 16);

  const char* bal_env = getenv("BHIP_COLOUR_BALANCE");
  const bool balance = !(bal_env && bal_env[0] == '0');
  for (int b = 0; b < L.nblocks; ++b) {
    const int n0 = L.node_start[b], n1 = L.node_start[b + 1];
    const int nown = n1 - n0;
    belems.clear();
    for (int i = n0; i < n1; ++i)
      for (int64_t a = adj_off[i]; a < adj_off[i + 1]; ++a) {
        int e = adj[a];
        if (elem_stamp[e] != b) {
          elem_stamp[e] = b;
          belems.push_back(e);
        }
      }
    // local ids: owned first (internal order), halo after in first-seen order
    // (whole columns at a time for column-numbered prisms, plane by plane)
    int nloc = nown;
    for (int e : belems)
      for (int k = 0; k < ndp; ++k) {
        int g = L.iperm[ikle[(long)k * nelem + e]];
        if (g >= n0 && g < n1) continue;
        if (node_stamp[g] != b) {
          if (L.col_ids) {
            const int c0 = g - g % nplan;
            for (int q = c0; q < c0 + nplan; ++q) {
              node_stamp[q] = b;
              node_local[q] = nloc++;
              L.halo.push_back(q);
            }
          } else {
            node_stamp[g] = b;
            node_local[g] = nloc++;
            L.halo.push_back(g);
          }
        }
      }
    if (nloc > 65535) return "block local node count exceeds 16-bit ids";
    L.halo_start.push_back((int32_t)L.halo.size());
    L.max_local = std::max(L.max_local, nloc);
    L.max_owned = std::max(L.max_owned, nown);
    auto local_of = [&](int g) { return (g >= n0 && g < n1) ? g - n0 : node_local[g]; };
    // element order inside the block: by smallest local id (LDS locality)
    std::sort(belems.begin(), belems.end(), [&](int a, int c) {
      int ma = 1 << 30, mc = 1 << 30;
      for (int k = 0; k < ndp; ++k) {
        ma = std::min(ma, local_of(L.iperm[ikle[(long)k * nelem + a]]));
        mc = std::min(mc, local_of(L.iperm[ikle[(long)k * nelem + c]]));
      }
      return ma != mc ? ma < mc : a < c;
    });
    // greedy colouring: conflicts only through owned nodes
    owned_mask.assign(nown, 0);
    bcolor.resize(belems.size());
    int ncol = 0;
    for (size_t s = 0; s < belems.size(); ++s) {
      int e = belems[s];
      uint64_t used = 0;
      for (int k = 0; k < ndp; ++k) {
        int l = local_of(L.iperm[ikle[(long)k * nelem + e]]);
        if (l < nown) used |= owned_mask[l];
      }
      int c = 0;
      while (c < 64 && (used >> c & 1)) ++c;
      if (c >= kMaxColors) return "block needs more colours than supported";
      bcolor[s] = c;
      ncol = std::max(ncol, c + 1);
      for (int k = 0; k < ndp; ++k) {
        int l = local_of(L.iperm[ikle[(long)k * nelem + e]]);
        if (l < nown) owned_mask[l] |= (1ull << c);
      }
    }
    // balance: first-fit fills the first colours beyond one slot per thread
    // of the workgroup, and a phase's slots past that count are loaded
    // outside the kernel's prefetch. Move elements out of such colours into
    // a colour with room that none of their owned nodes uses (each owned
    // node has at most one element per colour, so its mask bit is that
    // element's alone). BHIP_COLOUR_BALANCE=0 keeps first-fit as it is.
    if (balance) {
      std::vector<int> ccount(ncol, 0);
      for (size_t s = 0; s < belems.size(); ++s) ccount[bcolor[s]]++;
      for (int pass = 0; pass < 4; ++pass) {
        bool over = false;
        for (size_t s = 0; s < belems.size(); ++s) {
          const int o = bcolor[s];
          if (ccount[o] <= kPhaseSlots) continue;
          const int e = belems[s];
          int lo[8];
          uint64_t used = 0;
          for (int k = 0; k < ndp; ++k) {
            lo[k] = local_of(L.iperm[ikle[(long)k * nelem + e]]);
            if (lo[k] < nown) used |= owned_mask[lo[k]];
          }
          int c = 0;
          while (c < ncol && (c == o || ccount[c] >= kPhaseSlots || (used >> c & 1))) ++c;
          if (c == ncol) {
            over = true;
            continue;
          }
          for (int k = 0; k < ndp; ++k)
            if (lo[k] < nown) owned_mask[lo[k]] = (owned_mask[lo[k]] & ~(1ull << o)) | (1ull << c);
          bcolor[s] = c;
          ccount[o]--;
          ccount[c]++;
        }
        if (!over) break;
      }
    }
    // slots sorted by colour (stable: keeps the locality order inside)
    sorted.resize(belems.size());
    std::iota(sorted.begin(), sorted.end(), 0);
    std::stable_sort(sorted.begin(), sorted.end(),
                     [&](int a, int c) { return bcolor[a] < bcolor[c]; });
    int32_t* co = &L.color_off[(size_t)b * (kMaxColors + 1)];
    const int slot0 = (int)slot_elem.size();
    for (int c = 0; c <= kMaxColors; ++c) co[c] = slot0;
    for (size_t q = 0; q < sorted.size(); ++q) {
      int s = sorted[q];
      int e = belems[s];
      slot_elem.push_back(e);
      for (int k = 0; k < ndp; ++k)
        lid_cols[k].push_back((uint16_t)local_of(L.iperm[ikle[(long)k * nelem + e]]));
    }
    // colour c occupies [co[c], co[c+1])
    {
      std::vector<int> cnt(kMaxColors + 1, 0);
      for (size_t s = 0; s < belems.size(); ++s) cnt[bcolor[s]]++;
      int acc = slot0;
      for (int c = 0; c < kMaxColors; ++c) {
        co[c] = acc;
        acc += cnt[c];
      }
      co[kMaxColors] = acc;
      for (int c = ncol; c < kMaxColors; ++c) co[c] = acc;
    }
    L.ncolor[b] = ncol;
    L.max_colors = std::max(L.max_colors, ncol);
  }
  L.halo_total = (int64_t)L.halo.size();
  L.nslot = (int)slot_elem.size();
  L.nslot_pad = (L.nslot + 255) / 256 * 256;
  L.slot_elem = std::move(slot_elem);
  L.lid.assign((size_t)ndp * L.nslot_pad, 0);
  for (int k = 0; k < ndp; ++k)
    std::copy(lid_cols[k].begin(), lid_cols[k].end(),
              L.lid.begin() + (size_t)k * L.nslot_pad);
  return validate_layout(L, ikle);
}

std::string validate_layout(const Layout& L, const int32_t* ikle) {
  const int ndp = L.ndp, nelem = L.nelem, npoin = L.npoin;
  if ((int)L.perm.size() != npoin || (int)L.iperm.size() != npoin) return "perm size";
  for (int i = 0; i < npoin; ++i) {
    if (L.perm[i] < 0 || L.perm[i] >= npoin) return "perm range";
    if (L.iperm[L.perm[i]] != i) return "perm not a bijection";
  }
  if (L.node_start[0] != 0 || L.node_start[L.nblocks] != npoin) return "node ranges";
  std::vector<int32_t> block_of(npoin, -1);
  for (int b = 0; b < L.nblocks; ++b) {
    if (L.node_start[b + 1] < L.node_start[b]) return "node ranges not monotone";
    for (int g = L.node_start[b]; g < L.node_start[b + 1]; ++g) block_of[g] = b;
  }
  // every element appears in every block owning one of its nodes, once
  std::vector<int32_t> seen_count(nelem, 0);
  std::vector<int32_t> need(nelem, 0);
  for (int e = 0; e < nelem; ++e) {
    int blocks[6];
    int nbk = 0;
    for (int k = 0; k < ndp; ++k) {
      int g = L.iperm[ikle[(long)k * nelem + e]];
      int bb = block_of[g];
      bool dup = false;
      for (int q = 0; q < nbk; ++q) dup |= (blocks[q] == bb);
      if (!dup) blocks[nbk++] = bb;
    }
    need[e] = nbk;
  }
  std::vector<int32_t> mark(npoin, -1);
  for (int b = 0; b < L.nblocks; ++b) {
    const int n0 = L.node_start[b], n1 = L.node_start[b + 1], nown = n1 - n0;
    const int h0 = L.halo_start[b], h1 = L.halo_start[b + 1];
    const int nloc = nown + (h1 - h0);
    if (nloc > L.max_local || nown > L.max_owned) return "block size bounds";
    const int32_t* co = &L.color_off[(size_t)b * (kMaxColors + 1)];
    if (L.ncolor[b] > kMaxColors) return "colour count";
    for (int c = 0; c < kMaxColors; ++c)
      if (co[c] > co[c + 1]) return "colour offsets not monotone";
    if (co[kMaxColors] > L.nslot || co[0] < 0) return "colour offsets out of range";
    if (b + 1 < L.nblocks && co[kMaxColors] != L.color_off[(size_t)(b + 1) * (kMaxColors + 1)])
      return "blocks not contiguous in slots";
    for (int c = 0; c < kMaxColors; ++c) {
      for (int s = co[c]; s < co[c + 1]; ++s) {
        int e = L.slot_elem[s];
        if (e < 0 || e >= nelem) return "slot element out of range";
        seen_count[e]++;
        for (int k = 0; k < ndp; ++k) {
          int l = L.lid[(size_t)k * L.nslot_pad + s];
          if (l >= nloc) return "local id out of block range";
          int g = (l < nown) ? n0 + l : L.halo[h0 + l - nown];
          if (L.perm[g] != ikle[(long)k * nelem + e]) return "local id maps to the wrong node";
          if (L.col_ids && k >= 3 && l != L.lid[(size_t)(k - 3) * L.nslot_pad + s] + 1)
            return "column ids: top node is not the bottom node's successor";
          if (l < nown) {
            // conflict check: an owned node at most once per colour
            int tag = b * kMaxColors + c;
            if (mark[g] == tag) return "colour conflict on an owned node";
            mark[g] = tag;
          }
        }
      }
    }
  }
  for (int e = 0; e < nelem; ++e)
    if (seen_count[e] != need[e]) return "element instance count mismatch";
  return "";
}

// Layout of an ensemble of independent meshes: each mesh is laid out on its
// own (its blocks hold only its nodes) and the layouts are concatenated, so
// one launch covers every mesh. Node m of mesh s is node noff[s] + m of the
// union, element e is element eoff[s] + e.
std::string build_layout_multi(int nsys, const int* npoin, const int* nelem, int ndp,
                               const int32_t* ikle, const double* x, const double* y,
                               int block_nodes, Layout& L, std::vector<int32_t>& sys_blocks,
                               std::vector<int32_t>& ikle_union) {
  if (nsys < 1) return "no mesh";
  std::vector<Layout> parts(nsys);
  std::vector<int64_t> noff(nsys + 1, 0), eoff(nsys + 1, 0);
  for (int s = 0; s < nsys; ++s) {
    noff[s + 1] = noff[s] + npoin[s];
    eoff[s + 1] = eoff[s] + nelem[s];
  }
  if (noff[nsys] > INT32_MAX || eoff[nsys] * ndp > INT32_MAX) return "ensemble too large";
  const int NP = (int)noff[nsys], NE = (int)eoff[nsys];
  ikle_union.assign((size_t)ndp * NE, 0);
  for (int s = 0; s < nsys; ++s) {
    const int32_t* ik = ikle + (size_t)ndp * eoff[s];
    std::string err = build_layout(npoin[s], nelem[s], ndp, ik, x ? x + noff[s] : nullptr,
                                   y ? y + noff[s] : nullptr, 1, block_nodes, parts[s]);
    if (!err.empty()) return "mesh " + std::to_string(s) + ": " + err;
    for (int k = 0; k < ndp; ++k)
      for (int e = 0; e < nelem[s]; ++e)
        ikle_union[(size_t)k * NE + eoff[s] + e] = (int32_t)(ik[(size_t)k * nelem[s] + e] + noff[s]);
  }
  L = Layout();
  L.npoin = NP;
  L.nelem = NE;
  L.ndp = ndp;
  L.nplan = 1;
  L.block_nodes = parts[0].block_nodes;
  L.perm.resize(NP);
  L.iperm.resize(NP);
  sys_blocks.assign(nsys + 1, 0);
  int nslot = 0;
  for (int s = 0; s < nsys; ++s) nslot += parts[s].nslot;
  L.nslot = nslot;
  L.nslot_pad = (nslot + 255) / 256 * 256;
  L.lid.assign((size_t)ndp * L.nslot_pad, 0);
  L.slot_elem.resize(nslot);
  L.node_start.push_back(0);
  L.halo_start.push_back(0);
  int so = 0;
  for (int s = 0; s < nsys; ++s) {
    const Layout& P = parts[s];
    const int no = (int)noff[s], eo = (int)eoff[s], ho = (int)L.halo.size();
    for (int i = 0; i < P.npoin; ++i) {
      L.perm[no + i] = no + P.perm[i];
      L.iperm[no + i] = no + P.iperm[i];
    }
    for (int b = 0; b < P.nblocks; ++b) {
      L.node_start.push_back(no + P.node_start[b + 1]);
      L.halo_start.push_back(ho + P.halo_start[b + 1]);
      for (int c = 0; c <= kMaxColors; ++c)
        L.color_off.push_back(so + P.color_off[(size_t)b * (kMaxColors + 1) + c]);
      L.ncolor.push_back(P.ncolor[b]);
    }
    for (int32_t h : P.halo) L.halo.push_back(no + h);
    for (int k = 0; k < ndp; ++k)
      std::copy(P.lid.begin() + (size_t)k * P.nslot_pad,
                P.lid.begin() + (size_t)k * P.nslot_pad + P.nslot,
                L.lid.begin() + (size_t)k * L.nslot_pad + so);
    for (int q = 0; q < P.nslot; ++q) L.slot_elem[so + q] = eo + P.slot_elem[q];
    sys_blocks[s + 1] = sys_blocks[s] + P.nblocks;
    L.max_colors = std::max(L.max_colors, P.max_colors);
    L.max_local = std::max(L.max_local, P.max_local);
    L.max_owned = std::max(L.max_owned, P.max_owned);
    so += P.nslot;
  }
  L.nblocks = sys_blocks[nsys];
  L.halo_total = (int64_t)L.halo.size();
  return validate_layout(L, ikle_union.data());
}

}  // namespace bhip

