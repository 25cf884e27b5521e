// Wavefront tile K-way merge: every 64-lane wavefront is an independent
// merger with its own LDS windows and merge buffers, so a merge step never
// waits on a workgroup barrier and 12+ mergers share a CU to hide LDS and
// HBM latency. Same contract as k_tmerge (elements between two consistent
// cuts from k_partition, written at rank sum(lo)); same step:
//   refill windows -> frontier y (smallest last element over full windows) ->
//   gallop if a single list holds everything <= y, else gather the K taken
//   prefixes and merge them with log2 K stable merge-path passes -> store.
// Differences: lane a owns active list a (empty lists are compacted away,
// keeping list order for stability), and the window per list is
// WIN / pow2ceil(active lists), so a merger with few active lists (the last
// round, presorted input) gets long windows instead of short steps.
#pragma once
#include "gmms_elem.hxx"
#include "gmms_merge.hxx"

namespace gmms {

constexpr int WM_WAVES = 4;  // mergers per workgroup

// LDS words per wavefront merger with a window budget of WIN elements.
__host__ __device__ inline size_t wmerge_lds_words_per_wave(int W, int WIN) {
  return 3 * (size_t)WIN * W + 4 * WAVE + 8;
}

__device__ __forceinline__ void wave_lds_sync() {
  // LDS instructions of one wavefront execute in order; this keeps the
  // compiler from moving LDS accesses across phase boundaries.
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <class E>
__device__ __forceinline__ E wshfl(const E& e, int src) {
  uint32_t w[E::W];
#pragma unroll
  for (int q = 0; q < E::W; ++q) w[q] = (uint32_t)__shfl((int)word(e, q), src, WAVE);
  return from_words<E>(w);
}
template <class E>
__device__ __forceinline__ E wxor(const E& e, int m) {
  uint32_t w[E::W];
#pragma unroll
  for (int q = 0; q < E::W; ++q) w[q] = (uint32_t)__shfl_xor((int)word(e, q), m, WAVE);
  return from_words<E>(w);
}
template <class E>
__device__ __forceinline__ E lget(const uint32_t* base, int cap, int idx) {
  uint32_t w[E::W];
#pragma unroll
  for (int q = 0; q < E::W; ++q) w[q] = base[q * cap + idx];
  return from_words<E>(w);
}
template <class E>
__device__ __forceinline__ void lput(uint32_t* base, int cap, int idx, const E& e) {
#pragma unroll
  for (int q = 0; q < E::W; ++q) base[q * cap + idx] = word(e, q);
}

// One stable pairwise merge pass: groups of g runs (run r at s_off[r]) are
// merged in halves; lane l writes outputs [l*Eo, (l+1)*Eo). FROM_WIN: the
// runs are list prefixes in the circular windows (run r = window r from
// s_head[r]) instead of being packed in src. WM_CAP: the word-plane stride.
// Reads one past a run's end stay inside this merger's LDS (index <= WM_CAP)
// and are never selected.
template <class E, bool FROM_WIN, int WM_CAP>
__device__ __forceinline__ void merge_pass(const uint32_t* src, uint32_t* dst, int g, int Kp, int S, int Eo, int lane,
                                           const int* s_off, const int* s_head, int F) {
  int o = lane * Eo;
  const int oend = min(S, o + Eo);
  while (o < oend) {
    int l = 0, h = Kp / g - 1;
    while (l < h) {
      const int mid = (l + h + 1) >> 1;
      const bool c = s_off[mid * g] <= o;
      l = c ? mid : l;
      h = c ? h : mid - 1;
    }
    const int gs = s_off[l * g], gm = s_off[l * g + g / 2], ge = s_off[l * g + g];
    const int n1 = gm - gs, n2 = ge - gm;
    const int d = o - gs;
    // element addresses of the two halves
    int b1, b2, h1 = 0, h2 = 0;
    if (FROM_WIN) {
      b1 = (l * g) * F;  // g == 2: runs 2l and 2l+1
      b2 = b1 + F;
      h1 = s_head[l * g];
      h2 = s_head[l * g + 1];
    } else {
      b1 = gs;
      b2 = gm;
    }
    const int msk = F - 1;
    auto adr1 = [&](int i) { return FROM_WIN ? b1 + ((h1 + i) & msk) : b1 + i; };
    auto adr2 = [&](int i) { return FROM_WIN ? b2 + ((h2 + i) & msk) : b2 + i; };
    int plo = d - n2 > 0 ? d - n2 : 0;
    int phi = d < n1 ? d : n1;
    while (plo < phi) {
      const int mid = (plo + phi) >> 1;
      // A1[mid] is among the first d outputs iff A1[mid] <= A2[d-1-mid] (left run first on ties)
      const bool c = !lt(lget<E>(src, WM_CAP, adr2(d - 1 - mid)), lget<E>(src, WM_CAP, adr1(mid)));
      plo = c ? mid + 1 : plo;
      phi = c ? phi : mid;
    }
    int i = plo, jj = d - plo;
    const int cntm = min(oend, ge) - o;
    E va = lget<E>(src, WM_CAP, adr1(i));
    E vb = lget<E>(src, WM_CAP, adr2(jj));
    for (int e = 0; e < cntm; ++e) {
      const bool ta = (jj >= n2) | ((i < n1) & !lt(vb, va));
      lput<E>(dst, WM_CAP, o + e, ta ? va : vb);
      i += ta ? 1 : 0;
      jj += ta ? 0 : 1;
      const E nv = lget<E>(src, WM_CAP, ta ? adr1(i) : adr2(jj));
      va = ta ? nv : va;
      vb = ta ? vb : nv;
    }
    o += cntm;
  }
}

// One merger, executed by one wavefront using the LDS region `my`.
template <int DT, bool CONV_OUT, int WIN>
__device__ __forceinline__ void wmerge_one(const RoundArgs& a, const MergerInfo& mi, int lane, uint32_t* my) {
  using E = typename DType<DT>::E;
  constexpr int W = E::W;
  constexpr int CHUNKS = WIN / WAVE;
  uint32_t* win = my;                 // [list a][F] circular windows, W planes of WIN
  uint32_t* bufA = win + WIN * W;
  uint32_t* bufB = bufA + WIN * W;
  int* st = (int*)(bufB + WIN * W);   // per active list: head, taken, off (+1)
  int* s_head = st;
  int* s_take = st + WAVE;
  int* s_off = st + 2 * WAVE;         // WAVE + 1 entries

  const int K = mi.K;
  const int j = mi.j;
  const int64_t tbase = mi.tbase;

  // lane k: list k's range [lo, hi) for this merger
  int64_t lo = 0, hi = 0;
  if (lane < K) {
    const int64_t nk = mlist_len(mi, lane);
    lo = j == 0 ? 0 : a.cuts[(mi.cut_row + j) * a.cut_stride + lane];
    hi = j == mi.m - 1 ? nk : a.cuts[(mi.cut_row + j + 1) * a.cut_stride + lane];
  }
  int64_t obase = tbase + wave_sum64(lo);
  // compact the non-empty lists, keeping list order
  const bool act = lane < K && hi > lo;
  const unsigned long long amask = __ballot(act);
  const int Ka = __popcll(amask);
  if (Ka == 0) return;
  const int myidx = __popcll(amask & ((1ull << lane) - 1ull));
  // lane a takes the state of the a-th active list
  int src_lane = 0;
  {
    // find the lane whose rank among active lanes is `lane`
    const int sel = act ? myidx : -1;
    for (int l = 0; l < WAVE; ++l) {
      const int sl = __shfl(sel, l, WAVE);
      if (sl == lane) src_lane = l;
    }
  }
  const int64_t lb = tbase + (int64_t)src_lane * mi.L;
  int64_t cur = lb + __shfl(lo, src_lane, WAVE);
  int64_t endp = lb + __shfl(hi, src_lane, WAVE);
  const bool live = lane < Ka;
  if (!live) cur = endp = 0;
  int Kp = 1;
  while (Kp < Ka) Kp <<= 1;
  const int F = WIN / Kp;  // window per active list (power of two >= WIN/64)
  int head = 0, cnt = 0;

  for (;;) {
    const int64_t rem = endp - cur;
    const int want = live ? (int)(rem < F ? rem : F) : 0;
    if (__ballot(want > 0) == 0) break;

    // ---- refill: chunk u covers window slots u*64 .. u*64+63
    if (lane < Ka) s_head[lane] = head;
    wave_lds_sync();
    {
      // all loads in flight first; remember which slots got one
      E v[CHUNKS];
      uint32_t vmask = 0;
      // Loads are unconditional (a slot with nothing to fetch reads element 0,
      // always in bounds) so that no branch join forces a wait between them.
      if (F >= WAVE) {
        // a chunk lies inside one list: its state is a wave-uniform readlane
#pragma unroll
        for (int u = 0; u < CHUNKS; ++u) {
          const int slot = u * WAVE + lane;
          const int lr = (u * WAVE) / F;
          const int l = lr < WAVE ? lr : WAVE - 1;
          const int sj = slot & (F - 1);
          const int hl = __builtin_amdgcn_readlane(head, l);
          const int cl = __builtin_amdgcn_readlane(cnt, l);
          const int wl = __builtin_amdgcn_readlane(want, l);
          const int64_t cu = (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(cur >> 32), l) << 32) |
                                       (uint32_t)__builtin_amdgcn_readlane((int)cur, l));
          const int i = (sj - hl) & (F - 1);
          const bool ok = lr < Ka && i >= cl && i < wl;
          v[u] = gload<DT, false>(a.in_keys, a.in_vals, ok ? cu + i : 0);
          vmask |= ok ? (1u << u) : 0u;
        }
      } else {
#pragma unroll
        for (int u = 0; u < CHUNKS; ++u) {
          const int slot = u * WAVE + lane;
          const int l = slot / F;  // active list of this slot (< 64: F >= 16)
          const int sj = slot & (F - 1);
          const int hl = __shfl(head, l, WAVE);
          const int cl = __shfl(cnt, l, WAVE);
          const int wl = __shfl(want, l, WAVE);
          const int64_t cu = __shfl(cur, l, WAVE);
          const int i = (sj - hl) & (F - 1);  // logical index of this slot
          const bool ok = l < Ka && i >= cl && i < wl;
          v[u] = gload<DT, false>(a.in_keys, a.in_vals, ok ? cu + i : 0);
          vmask |= ok ? (1u << u) : 0u;
        }
      }
#pragma unroll
      for (int u = 0; u < CHUNKS; ++u) {
        if ((vmask >> u) & 1u) lput<E>(win, WIN, u * WAVE + lane, v[u]);
      }
    }
    wave_lds_sync();

    // ---- frontier y = min over full windows of the last element, tag (elem, list)
    const bool full = live && want == F;
    E y = full ? lget<E>(win, WIN, lane * F + ((head + F - 1) & (F - 1))) : Limits<E>::hi();
    int ky = full ? lane : WAVE;
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
      const E oy = wxor(y, d);
      const int oky = __shfl_xor(ky, d, WAVE);
      const bool take = oky < WAVE && (ky == WAVE || lt(oy, y) || (!lt(y, oy) && oky < ky));
      y = take ? oy : y;
      ky = take ? oky : ky;
    }
    int sk;
    bool gallop = false;
    int64_t glen = 0;
    if (ky == WAVE) {
      sk = want;  // everything left is in the windows
    } else {
      const bool ne = live && want > 0 && lane != ky;
      E z = ne ? lget<E>(win, WIN, lane * F + head) : Limits<E>::hi();
      int kz = ne ? lane : WAVE;
#pragma unroll
      for (int d = 1; d < WAVE; d <<= 1) {
        const E oz = wxor(z, d);
        const int okz = __shfl_xor(kz, d, WAVE);
        const bool take = okz < WAVE && (kz == WAVE || lt(oz, z) || (!lt(z, oz) && okz < kz));
        z = take ? oz : z;
        kz = take ? okz : kz;
      }
      gallop = kz == WAVE || lt(y, z) || (!lt(z, y) && ky < kz);
      if (gallop) {
        // run of list ky (from cur + F) that precedes z
        const int64_t base = __shfl(cur, ky, WAVE);
        int64_t glo = F, ghi = __shfl(endp, ky, WAVE) - base;
        if (kz != WAVE) {
          while (ghi > glo) {
            const int64_t span = ghi - glo;
            const int64_t step = (span + WAVE - 1) / WAVE;
            const int64_t p = glo + (int64_t)lane * step;
            bool before = true;
            if (p < ghi) {
              const E v = gload<DT, false>(a.in_keys, a.in_vals, base + p);
              before = lt(v, z) || (!lt(z, v) && ky < kz);
            }
            const unsigned long long nb = __ballot(!before && p < ghi);
            if (nb == 0) {
              const int64_t np = (ghi - 1 - glo) / step;
              glo = glo + (np < WAVE - 1 ? np : WAVE - 1) * step + 1;
            } else {
              const int l0 = __ffsll((long long)nb) - 1;
              const int64_t phi = glo + (int64_t)l0 * step;
              glo = l0 == 0 ? glo : glo + (int64_t)(l0 - 1) * step + 1;
              ghi = phi;
            }
          }
        } else {
          glo = ghi;
        }
        glen = glo;
        sk = 0;
      } else {
        sk = 0;
        if (live && want > 0) {
          if (lane == ky) {
            sk = F;
          } else {
            int l = 0, h = want;
            while (l < h) {
              const int mid = (l + h) >> 1;
              const E v = lget<E>(win, WIN, lane * F + ((head + mid) & (F - 1)));
              const bool tk = lane < ky ? !lt(y, v) : lt(v, y);
              l = tk ? mid + 1 : l;
              h = tk ? h : mid;
            }
            sk = l;
          }
        }
      }
    }

    if (gallop) {
      const int64_t src = __shfl(cur, ky, WAVE);
      for (int64_t i = lane; i < glen; i += WAVE)
        gstore<DT, CONV_OUT>(a.out_keys, a.out_vals, obase + i, gload<DT, false>(a.in_keys, a.in_vals, src + i));
      if (lane == ky) {
        cur += glen;
        head = 0;
        cnt = 0;
      }
      obase += glen;
      wave_lds_sync();
      continue;
    }

    // ---- offsets of the taken prefixes
    int incl = sk;
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
      const int o = __shfl_up(incl, d, WAVE);
      incl += lane >= d ? o : 0;
    }
    const int S = __shfl(incl, WAVE - 1, WAVE);
    s_take[lane] = sk;
    s_off[lane] = incl - sk;
    if (lane == WAVE - 1) s_off[WAVE] = S;
    wave_lds_sync();

    // ---- stable pairwise merge passes over Kp runs at s_off[]; the first
    // pass reads the runs straight out of the circular windows
    const int Eo = (S + WAVE - 1) / WAVE;
    if (Kp == 1) {
      // one run: it is window 0 (list 0's, held by lane 0), in order from its head
      const int h0 = __builtin_amdgcn_readlane(head, 0);
      for (int o = lane; o < S; o += WAVE)
        gstore<DT, CONV_OUT>(a.out_keys, a.out_vals, obase + o, lget<E>(win, WIN, (h0 + o) & (F - 1)));
    } else {
      // gather the taken prefixes into bufA (run r at s_off[r])
#pragma unroll
      for (int u = 0; u < CHUNKS; ++u) {
        const int slot = u * WAVE + lane;
        const int l = slot / F;
        const int i = slot & (F - 1);  // logical index within list l's prefix
        if (l < Ka) {
          const int tl = s_take[l];
          if (i < tl) lput<E>(bufA, WIN, s_off[l] + i, lget<E>(win, WIN, l * F + ((s_head[l] + i) & (F - 1))));
        }
      }
      wave_lds_sync();
      uint32_t* src = bufA;
      uint32_t* dst = bufB;
      for (int g = 2; g <= Kp; g <<= 1) {
        merge_pass<E, false, WIN>(src, dst, g, Kp, S, Eo, lane, s_off, s_head, F);
        wave_lds_sync();
        uint32_t* t = src;
        src = dst;
        dst = t;
      }
      for (int o = lane; o < S; o += WAVE) gstore<DT, CONV_OUT>(a.out_keys, a.out_vals, obase + o, lget<E>(src, WIN, o));
    }

    // ---- advance
    obase += S;
    cur += sk;
    head = (head + sk) & (F - 1);
    cnt = want - sk;
    wave_lds_sync();
  }
}

// Plain sort: one merger per wavefront over the uniform task grid.
template <int DT, bool CONV_OUT, int WIN>
__global__ __launch_bounds__(WM_WAVES* WAVE) void k_wmerge(RoundArgs a) {
  using E = typename DType<DT>::E;
  extern __shared__ uint32_t lds[];
  const int lane = threadIdx.x & (WAVE - 1);
  const int wv = threadIdx.x / WAVE;
  const int64_t merger = (int64_t)blockIdx.x * WM_WAVES + wv;
  if (merger >= a.tasks * a.m) return;  // whole wavefront leaves; no barriers in this kernel
  uint32_t* my = lds + (size_t)wv * wmerge_lds_words_per_wave(E::W, WIN);
  wmerge_one<DT, CONV_OUT, WIN>(a, plain_merger(a, merger), lane, my);
}

// Batched segments: a persistent launch; every wavefront takes merger
// descriptors from a.items through the a.queue counter until the list is
// exhausted (every wavefront reaches that exit). The counter starts at zero
// (cleared by the launch before).
template <int DT, bool CONV_OUT, int WIN>
__global__ __launch_bounds__(WM_WAVES* WAVE) void k_wmerge_persistent(RoundArgs a) {
  using E = typename DType<DT>::E;
  extern __shared__ uint32_t lds[];
  const int lane = threadIdx.x & (WAVE - 1);
  const int wv = threadIdx.x / WAVE;
  uint32_t* my = lds + (size_t)wv * wmerge_lds_words_per_wave(E::W, WIN);
  for (;;) {
    unsigned int it = 0;
    if (lane == 0) it = atomicAdd(a.queue, 1u);
    it = (unsigned int)__shfl((int)it, 0, WAVE);
    if ((int64_t)it >= a.n_items) break;
    wmerge_one<DT, CONV_OUT, WIN>(a, load_merger(a.items + it), lane, my);
    wave_lds_sync();
  }
}

}  // namespace gmms
