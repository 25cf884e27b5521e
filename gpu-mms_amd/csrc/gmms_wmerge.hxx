// Wavefront tile K-way merge: every 64-lane wavefront is an independent
// merger with its own LDS windows, so a merge step never waits on a
// workgroup barrier and 16 mergers share a CU to hide LDS and HBM latency.
// A merger merges the elements between two consistent cuts from k_partition
// and writes them at rank sum(lo). One step:
//   refill windows -> frontier y (smallest last element over full windows) ->
//   gallop if a single list holds everything <= y, else take every list's
//   prefix <= y into a register tile and merge the runs with bitonic merge
//   levels -> store.
// Lane a owns active list a (empty lists are compacted away), and the window
// per list is F = WIN / pow2ceil(active lists), so a merger with few active
// lists (the last round, presorted input) gets long windows instead of short
// steps.
// The register tile has WIN slots, NREG = WIN/64 per lane (slot s in lane
// s/NREG, register s%NREG); run r fills slots [r*F, r*F + F): its taken
// prefix, then the maximum element. Each run is sorted, so the tile is Kp
// sorted blocks of F and the merge levels of the base-case network from
// block size F up to WIN sort it; the first S slots are the S taken elements
// in order. Filler slots hold the all-ones element, which equals the
// largest possible real element bit for bit, so a real maximum and a filler
// are interchangeable and the first S slots are exact.
#pragma once
#include <type_traits>
#include "gmms_basecase.hxx"
#include "gmms_elem.hxx"
#include "gmms_merge.hxx"

namespace gmms {

constexpr int WM_WAVES = 4;  // mergers per workgroup of the persistent (batched) launch
// The plain launch puts all of a CU's resident mergers in one workgroup
// (WM_PLAIN_MAX at most) so that the mergers sharing a SIMD can compare
// their progress through LDS.
constexpr int WM_PLAIN_MAX = 16;

// LDS words per wavefront merger with a window budget of WIN elements: the
// windows, and the transpose of the register tile for the store (multi-word
// elements: bank-padded rows; three-word pairs hold half the tile's lanes at
// a time, so that they too fit 16 mergers in a CU's LDS).
__host__ __device__ constexpr int wmerge_xpose_rows(int W) { return W >= 3 ? WAVE / 2 : WAVE; }
__host__ __device__ inline size_t wmerge_lds_words_per_wave(int W, int WIN) {
  return W == 1 ? (size_t)2 * WIN : (size_t)WIN * W + (size_t)wmerge_xpose_rows(W) * (WIN / WAVE + 1) * W;
}
// Workgroup LDS layout: the windows of all its mergers first (merger wv's W
// planes of WIN words at wv * W * WIN), then their transposes. Every window
// plane then starts at a multiple of 4 * WIN bytes, so a window of F <= WIN
// words starts at a multiple of 4F bytes and a wrapped slot offset can be
// OR-ed into its address.
__device__ __forceinline__ uint32_t* wmerge_windows(uint32_t* lds, int wv, int W, int WIN) {
  return lds + (size_t)wv * W * WIN;
}
__device__ __forceinline__ uint32_t* wmerge_transpose(uint32_t* lds, int nw, int wv, int W, int WIN) {
  return lds + (size_t)nw * W * WIN + (size_t)wv * (wmerge_lds_words_per_wave(W, WIN) - (size_t)W * WIN);
}
// LDS words after the nw mergers' regions of a plain launch: every
// merger's SIMD, then four progress words per SIMD.
__host__ __device__ inline size_t wmerge_peer_words(int nw) { return (size_t)nw + 16; }

// Issue priority from the progress of the mergers that share this
// wavefront's SIMD. A SIMD's instruction arbiter favours its older
// wavefronts among equal priorities, so equal mergers sharing a SIMD would
// otherwise finish tens of microseconds apart (measured: 76 us of a 710 us
// round at 2^28 int32), the last ones alone on the SIMD and unable to hide
// their latencies. The merger with the most outputs left gets the highest
// priority, the one with the fewest the lowest, so the SIMD's mergers move
// in step and finish together. Each SIMD has four progress words (outputs
// left; all ones where no merger runs); a merger writes its own every step
// and reads its SIMD's four without ordering, which a priority tolerates.
struct WavePeers {
  uint32_t grp;   // LDS byte address of the SIMD's four words
  uint32_t mine;  // and of this merger's
};

__device__ __forceinline__ void wave_lds_sync() {
  // LDS instructions of one wavefront execute in order; this keeps the
  // compiler from moving LDS accesses across phase boundaries.
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <class E>
__device__ __forceinline__ E lget(const uint32_t* base, int cap, int idx) {
  uint32_t w[E::W];
#pragma unroll
  for (int q = 0; q < E::W; ++q) w[q] = base[q * cap + idx];
  return from_words<E>(w);
}
// The same read at an LDS byte address (word plane q at ab + 4 * q * cap).
// Window slots are addressed as ((position * 4) & (4F - 1)) | (LDS address of
// the window), one add and one and-or per slot: every window starts at a
// multiple of 4F bytes (the wavefront regions are multiples of 8 KB apart).
typedef __attribute__((address_space(3))) const uint32_t* lds_cu32;
__device__ __forceinline__ uint32_t lds_addr(const uint32_t* p) {
  return (uint32_t)(size_t)(lds_cu32)p;
}
template <class E>
__device__ __forceinline__ E lget_b(int cap, uint32_t ab) {
  uint32_t w[E::W];
#pragma unroll
  for (int q = 0; q < E::W; ++q) w[q] = *(lds_cu32)(size_t)(ab + 4u * (uint32_t)(q * cap));
  return from_words<E>(w);
}
template <class E>
__device__ __forceinline__ void lput(uint32_t* base, int cap, int idx, const E& e) {
#pragma unroll
  for (int q = 0; q < E::W; ++q) base[q * cap + idx] = word(e, q);
}

// Wave-uniform minimum of (x, k) over lanes 0..Kp-1 in (element, lane) order;
// lanes without a candidate pass (max element, WAVE). k = WAVE: none.
// Partner for level D of a minimum reduction whose result is read from
// lane 0: after level D, lane 0 must hold the minimum of lanes 0 .. 2D-1.
// Level 4 reads lane l + 4 (mod 16) with the DPP rotation row_ror:12 (no DPP
// xor-4 pattern exists), so lane 0 combines lanes 0-3 with 4-7 and lane 8
// lanes 8-11 with 12-15 for level 8; the other levels use the xor exchange.
template <int D>
__device__ __forceinline__ uint32_t rot_lane(uint32_t x) {
  if constexpr (D == 4) return dpp_mov<0x12C>(x);  // row_ror:12: lane l reads lane (l + 4) & 15
  else return xor_lane<D>(x);
}

// Lanes without a candidate hold the maximum, so the levels inside a row
// (D <= 8: DPP, no LDS) run whatever Kp is; only the levels that go through
// the LDS unit (16, 32) are skipped for Kp <= 16 (one uniform branch, not a
// test per level).
template <class E>
__device__ __forceinline__ void frontier_min(E x, int k, int Kp, E& y, int& ky) {
  // butterfly over lane groups of 2..Kp with the DPP / swizzle exchanges;
  // lane 0 ends with the minimum
  auto lvl = [&](auto dc) {
    constexpr int D = decltype(dc)::value;
    {
      uint32_t w[E::W];
#pragma unroll
      for (int r = 0; r < E::W; ++r) w[r] = rot_lane<D>(word(x, r));
      const E ox = from_words<E>(w);
      const int ok = (int)rot_lane<D>((uint32_t)k);
      const bool take = ok < WAVE && (k == WAVE || lt(ox, x) || (!lt(x, ox) && ok < k));
      x = sel(take, ox, x);
      k = take ? ok : k;
    }
  };
  lvl(std::integral_constant<int, 1>{});
  lvl(std::integral_constant<int, 2>{});
  lvl(std::integral_constant<int, 4>{});
  lvl(std::integral_constant<int, 8>{});
  if (Kp > 16) {
    lvl(std::integral_constant<int, 16>{});
    if (Kp > 32) lvl(std::integral_constant<int, 32>{});
  }
  uint32_t w[E::W];
#pragma unroll
  for (int q = 0; q < E::W; ++q) w[q] = (uint32_t)__builtin_amdgcn_readfirstlane((int)word(x, q));
  y = from_words<E>(w);
  ky = __builtin_amdgcn_readfirstlane(k);
}
// 32-bit elements: a 32-bit minimum (one DPP-operand v_min per level), then
// the lowest candidate lane holding it by a ballot (no 64-bit compares: they
// cost about ten 32-bit instructions each on gfx950)
template <>
__device__ __forceinline__ void frontier_min<E32>(E32 x, int k, int Kp, E32& y, int& ky) {
  const bool cand = k != WAVE;
  uint32_t m = cand ? x.k : 0xFFFFFFFFu;
  auto lvl = [&](auto dc) {
    constexpr int D = decltype(dc)::value;
    const uint32_t o = rot_lane<D>(m);
    m = o < m ? o : m;
  };
  lvl(std::integral_constant<int, 1>{});
  lvl(std::integral_constant<int, 2>{});
  lvl(std::integral_constant<int, 4>{});
  lvl(std::integral_constant<int, 8>{});
  if (Kp > 16) {
    lvl(std::integral_constant<int, 16>{});
    if (Kp > 32) lvl(std::integral_constant<int, 32>{});
  }
  y.k = (uint32_t)__builtin_amdgcn_readfirstlane((int)m);
  const unsigned long long at = __ballot(cand && x.k == y.k);  // candidate lanes are k == lane
  ky = at ? __ffsll((long long)at) - 1 : WAVE;
}

// LDS DMA of element base + off into word planes dst[q * plane] (base and
// dst wave-uniform, off a per-lane 32-bit offset so the loads use the
// scalar-base + vector-offset form; lane l's words land at dst[q * plane + l]).
template <int DT>
__device__ __forceinline__ void dma_elem(const RoundArgs& a, int64_t base, uint32_t off, uint32_t* dst, int plane) {
  using E = typename DType<DT>::E;
  typedef __attribute__((address_space(3))) void* lds_ptr;
  // byte offsets as 32-bit values, zero-extended: the address is then
  // scalar base + 32-bit vector offset, with no 64-bit vector arithmetic
  if constexpr (E::W == 1) {
    const char* kb = (const char*)((const uint32_t*)a.in_keys + base);
    __builtin_amdgcn_global_load_lds((const void*)(kb + (uint32_t)(off * 4u)), (lds_ptr)dst, 4, 0, 0);
  } else {
    const char* kb = (const char*)((const uint64_t*)a.in_keys + base);
    const char* kw = kb + (uint32_t)(off * 8u);
    __builtin_amdgcn_global_load_lds((const void*)kw, (lds_ptr)dst, 4, 0, 0);
    __builtin_amdgcn_global_load_lds((const void*)(kw + 4), (lds_ptr)(dst + plane), 4, 0, 0);
    if constexpr (E::W == 3) {
      const char* vb = (const char*)(a.in_vals + base);
      __builtin_amdgcn_global_load_lds((const void*)(vb + (uint32_t)(off * 4u)), (lds_ptr)(dst + 2 * plane), 4, 0,
                                       0);
    }
  }
}

// The same DMA with the key byte offset kb of a 32-bit-key element (4 bytes
// per element) or of a 64-bit key (8 bytes; the value of a pair at kb / 2).
template <int DT>
__device__ __forceinline__ void dma_elem_b(const RoundArgs& a, int64_t base, uint32_t kb, uint32_t* dst, int plane) {
  using E = typename DType<DT>::E;
  typedef __attribute__((address_space(3))) void* lds_ptr;
  if constexpr (E::W == 1) {
    const char* k0 = (const char*)((const uint32_t*)a.in_keys + base);
    __builtin_amdgcn_global_load_lds((const void*)(k0 + kb), (lds_ptr)dst, 4, 0, 0);
  } else {
    const char* kw = (const char*)((const uint64_t*)a.in_keys + base) + kb;
    __builtin_amdgcn_global_load_lds((const void*)kw, (lds_ptr)dst, 4, 0, 0);
    __builtin_amdgcn_global_load_lds((const void*)(kw + 4), (lds_ptr)(dst + plane), 4, 0, 0);
    if constexpr (E::W == 3) {
      const char* vb = (const char*)(a.in_vals + base) + (kb >> 1);
      __builtin_amdgcn_global_load_lds((const void*)vb, (lds_ptr)(dst + 2 * plane), 4, 0, 0);
    }
  }
}

// min(rem, F) for a non-negative 64-bit count, from its 32-bit halves: a
// 64-bit vector compare costs about ten 32-bit instructions on gfx950, and
// the empty asm keeps the compiler from folding the halves back into one.
__device__ __forceinline__ int clamp_count(int64_t rem, int F) {
  uint32_t hi = (uint32_t)((uint64_t)rem >> 32);
  const uint32_t lo = (uint32_t)rem;
  asm("" : "+v"(hi));
  const uint32_t m = lo < (uint32_t)F ? lo : (uint32_t)F;
  return hi ? F : (int)m;
}
// the same for a wave-uniform count (scalar instructions only)
__device__ __forceinline__ int clamp_count_u(int64_t rem, int F) {
  uint32_t hi = (uint32_t)((uint64_t)rem >> 32);
  const uint32_t lo = (uint32_t)rem;
  asm("" : "+s"(hi));
  const uint32_t m = lo < (uint32_t)F ? lo : (uint32_t)F;
  return hi ? F : (int)m;
}

// One merger, executed by one wavefront using the LDS region `my`.
// PEERS: issue priority by rank among the SIMD's mergers (wp); otherwise by
// the fraction of this merger's output written.
template <int DT, bool CONV_OUT, int WIN, bool PEERS>
__device__ __forceinline__ void wmerge_one(const RoundArgs& a, const MergerInfo& mi, int lane, uint32_t* win,
                                           uint32_t* xp, const WavePeers& wp) {
  using E = typename DType<DT>::E;
  constexpr int W = E::W;
  constexpr int CHUNKS = WIN / WAVE;
  constexpr int NREG = WIN / WAVE;        // register-tile slots per lane
  // Store layout. 32-bit keys (VEC): ov[4u + t] is output 256u + 4*lane + t,
  // stored 16 bytes per lane. Multi-word elements: ov[u] is output
  // u*64 + lane, stored one element per lane (the 16-byte form needs more
  // registers than those kernels have to spare).
  constexpr bool VEC = W == 1;
  constexpr int XC = NREG / 4;            // VEC: 16-byte chunks per lane = output rows of 256
  constexpr int XS = VEC ? NREG : NREG + 1;  // transpose row stride (padded for one-word reads)
  constexpr int XH = VEC ? 1 : WAVE / wmerge_xpose_rows(W);  // multi-word: transposed in XH parts
  constexpr int XW = VEC ? WAVE * XS : wmerge_xpose_rows(W) * XS;  // transpose words per plane
  static_assert(!VEC || XC == 4, "the 16-byte transpose covers 16 slots per lane");
  static_assert(!VEC || SAMP == 256, "a store row of 256 outputs holds one sample");
  // win: [list a][F] circular windows, W planes of WIN (wmerge_windows);
  // xp: transpose of the register tile, W planes of XW (wmerge_transpose)

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
  int64_t obase = uniform64(tbase + wave_sum64(lo));
  const int64_t o0 = obase;  // this merger's output is [o0, final obase)
  const int64_t total = uniform64(wave_sum64(hi - lo));  // its output count
  int prio = -1;                                          // issue priority set last
  // outputs written when the priority bands change: 8/16, 13/16, 15/16 of
  // the total, as 32-bit counts (a heuristic only: saturated for mergers of
  // 2^32 elements or more)
  const uint64_t tot32 = total < 0xFFFFFFFFll ? (uint64_t)total : 0xFFFFFFFFull;
  const uint32_t pb8 = PEERS ? 0u : (uint32_t)(tot32 / 2), pb13 = PEERS ? 0u : (uint32_t)(tot32 * 13 / 16),
                 pb15 = PEERS ? 0u : (uint32_t)(tot32 * 15 / 16);
  // compact the non-empty lists, keeping list order
  const bool act = lane < K && hi > lo;
  const unsigned long long amask = __ballot(act);
  const int Ka = __popcll(amask);
  if (Ka == 0) return;
  const uint32_t tot_left = (uint32_t)tot32;  // outputs left, saturated at 2^32 - 1
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
  const int fsh = __builtin_ctz((unsigned)F);
  const int tr = (lane * NREG) >> fsh;  // the run whose slots this lane holds in the register tile
  // key bytes per element in the DMA offsets, and this lane's share of them
  constexpr uint32_t KB = W == 1 ? 4u : 8u;
  uint32_t laneb = (uint32_t)lane * KB;
  asm("" : "+v"(laneb));  // opaque: keeps the refill offsets in key bytes (add, and, compare)
  // Window refill by LDS DMA: the logical range [c, w) of every active list
  // (physical slots (h + i) & (F-1) of its window) is loaded from global
  // memory straight into LDS; chunk u covers window slots u*64 .. u*64+63.
  // Completion is awaited with vmcnt(0) later, so the loads of the next
  // step overlap this step's merge network.
  auto refill = [&](int64_t cu_, int h_, int c_, int w_) {
    if (W == 1 && F == WAVE) {
      // 32-bit keys, 64-slot windows (16 lists: the common case): chunk u
      // is list u's window. Lane a packs ((h + c) & 63) * 4 | (w - c) * 4 << 12
      // and the byte offset of element cur + c; slot s of window u then
      // loads element byte jb = (4s - off) & 255 past it when jb < 4(w - c):
      // three readlanes, a subtract, an and and a compare per chunk, and few
      // scalar instructions (no per-chunk list index, skip test or field
      // decode). Lanes past the active lists carry w = c: nothing loads.
      const uint32_t pk =
          ((uint32_t)((h_ + c_) & (WAVE - 1)) * 4u) | ((uint32_t)(w_ - c_) * 4u << 12);
      // lane a's first element as a 64-bit address (one vector add for all
      // lists): its two readlanes are the chunk's scalar base as they are
      const uint64_t ep = (uint64_t)(size_t)a.in_keys + (uint64_t)(cu_ + c_) * 4u;
#pragma unroll
      for (int u = 0; u < CHUNKS; ++u) {
        const uint32_t ps = (uint32_t)__builtin_amdgcn_readlane((int)pk, u);
        const char* g = (const char*)(size_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(ep >> 32), u) << 32) |
                                              (uint32_t)__builtin_amdgcn_readlane((int)ep, u));
        const uint32_t jb = (laneb - (ps & 0xFFFu)) & (WAVE * 4u - 1u);
        typedef __attribute__((address_space(3))) void* lds_ptr;
        if (jb < (ps >> 12))
          __builtin_amdgcn_global_load_lds((const void*)(g + jb), (lds_ptr)(win + u * WAVE), 4, 0, 0);
      }
    } else if (F >= WAVE) {
      // a chunk lies inside one list: its state is a wave-uniform readlane
      // of (head | cnt << 10 | want << 21) and of the 64-bit position.
      // Slot u*64 + lane of list lr's window holds logical index
      // i = (so + lane - h) & (F-1), so = (u*64) & (F-1); with
      // j = (i - c) & (F-1) the slots to load are j < w - c, element
      // cur + c + j (i in [c, w) exactly when j < w - c, as 0 <= c <= w <= F):
      // one add, one and and one compare per chunk, in key bytes.
      const uint32_t pk = (uint32_t)h_ | ((uint32_t)c_ << 10) | ((uint32_t)w_ << 21);
      // recomputed per call (scalar instructions) rather than kept across
      // the loop, where they would be spilled and reloaded by readlanes
      int ka = Ka, fs = fsh;
      asm volatile("" : "+s"(ka), "+s"(fs));
#pragma unroll
      for (int u = 0; u < CHUNKS; ++u) {
        const int lr = (u * WAVE) >> fs;
        if (lr < ka) {
          const uint32_t ps = (uint32_t)__builtin_amdgcn_readlane((int)pk, lr);
          const int hl = (int)(ps & 1023u), cl = (int)((ps >> 10) & 2047u), wl = (int)(ps >> 21);
          if (cl < wl) {
            const int64_t cu = (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(cu_ >> 32), lr) << 32) |
                                         (uint32_t)__builtin_amdgcn_readlane((int)cu_, lr));
            const uint32_t sft = (uint32_t)(((u * WAVE) & ((1 << fs) - 1)) - hl - cl) * KB;
            const uint32_t jb = (laneb + sft) & ((KB << fs) - 1u);
            if (jb < (uint32_t)(wl - cl) * KB) dma_elem_b<DT>(a, cu + cl, jb, win + u * WAVE, WIN);
          }
        }
      }
    } else {
#pragma unroll
      for (int u = 0; u < CHUNKS; ++u) {
        const int slot = u * WAVE + lane;
        const int l = slot / F;  // active list of this slot (< 64: F >= 16)
        const int hl = __shfl(h_, l, WAVE);
        const int cl = __shfl(c_, l, WAVE);
        const int wl = __shfl(w_, l, WAVE);
        const int64_t cu = __shfl(cu_, l, WAVE);
        const int i = (slot - hl) & (F - 1);  // logical index of this slot
        if (l < Ka && i >= cl && i < wl) dma_elem<DT>(a, cu, (uint32_t)i, win + u * WAVE, WIN);
      }
    }
  };
  auto wait_dma = [&]() {
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): every DMA (and store) done
    wave_lds_sync();
  };

  // window state: list a's window holds its next `want` elements, logical
  // [0, want) at physical (head + i) & (F-1); cur = its first element
  int head = 0;
  int want = live ? clamp_count(endp - cur, F) : 0;
  refill(cur, head, 0, want);
  wait_dma();

  for (;;) {
    if (__ballot(want > 0) == 0) break;
    // Issue priority by progress: the SIMD's arbiter favours older
    // wavefronts, so mergers of equal size sharing a SIMD would otherwise
    // finish far apart (the last ones alone, unable to hide latency); the
    // merger furthest behind gets the highest priority. Plain launch: by
    // rank among the SIMD's mergers (WavePeers). Batched launch (mergers of
    // other workgroups share the SIMD): by the fraction written, in bands
    // that narrow towards the end (1/2, 5/16, 1/8, 1/16 of the output),
    // where the order among equal priorities decides who finishes last.
    {
      const uint32_t done = (uint32_t)obase - (uint32_t)o0;
      int lvl;
      if constexpr (PEERS) {
        // rank of this merger's outputs left among its SIMD's mergers:
        // lanes 0-3 read the four words, a ballot counts the smaller ones
        const uint32_t left = tot_left - done;
        // the addresses are rebuilt from scalars every step rather than held
        // in vector registers across the loop (pairs use all 128)
        uint32_t ga = wp.grp, ma = wp.mine;
        asm volatile("" : "+s"(ga), "+s"(ma));
        *(volatile __attribute__((address_space(3))) uint32_t*)(size_t)ma = left;
        const uint32_t o =
            *(const volatile __attribute__((address_space(3))) uint32_t*)(size_t)(ga + 4u * (uint32_t)(lane & 3));
        lvl = __popcll(__ballot(o < left) & 0xFull);
      } else {
        lvl = done < pb8 ? 3 : done < pb13 ? 2 : done < pb15 ? 1 : 0;
      }
      if (lvl != prio) {
        prio = lvl;
        if (lvl >= 2) {
          if (lvl == 3) __builtin_amdgcn_s_setprio(3);
          else __builtin_amdgcn_s_setprio(2);
        } else {
          if (lvl == 1) __builtin_amdgcn_s_setprio(1);
          else __builtin_amdgcn_s_setprio(0);
        }
      }
    }

    // the window state of this lane's tile run (see the register tile
    // below), fetched first so that the permutes' LDS round trip overlaps
    // the frontier and the searches
    const int wr = __shfl(want, tr, WAVE);
    const int hd = __shfl(head, tr, WAVE);

    // ---- frontier y = min over full windows of the last element, tag (elem, list)
    const bool full = live && want == F;
    E y;
    int ky;
    // this lane's window in bytes: lw4 | ((position * 4) & m4w)
    const uint32_t lw4 = lds_addr(win) + (uint32_t)(lane * F) * 4u, m4w = (uint32_t)F * 4u - 1u;
    uint32_t h4 = (uint32_t)head * 4u;
    asm("" : "+v"(h4));
    const E lastv = full ? lget_b<E>(WIN, ((h4 + m4w - 3u) & m4w) | lw4) : Limits<E>::hi();
    frontier_min<E>(lastv, full ? lane : WAVE, Kp, y, ky);
    // ---- elements each list takes: its window prefix <= y. Every element
    // still outside the windows is >= y (y is the smallest last element of a
    // full window), so these S elements are the next S outputs. Equal
    // elements are equal bit for bit, so their order across lists does not
    // matter: with many duplicates every list takes its copies of y at once
    // instead of one list per step.
    int sk = 0;
    bool gallop = false;
    int pre = 0;  // elements of one other list written ahead of a gallop
    if (ky == WAVE) {
      sk = want;  // no full window: everything left is in the windows
    } else {
      if (W == 1 && F == WAVE) {
        // 32-bit keys in 64-slot windows: binary lifting, six dependent reads and no loop
        // control. pos ends as the count of window elements <= y below 64;
        // a full window whose last element is <= y (it then equals y, the
        // smallest last element) takes all 64; a window holding fewer than
        // 64 elements (a list's last ones) bounds every step by its count
        // (taken only when some lane has such a window).
        if (live && want > 0) {
          uint32_t pos = 0;
          if (__ballot(want < F) == 0) {
#pragma unroll
            for (uint32_t b = WAVE / 2; b >= 1; b >>= 1) {
              const E v = lget_b<E>(WIN, ((h4 + 4u * (pos + b - 1u)) & m4w) | lw4);
              pos = lt(y, v) ? pos : pos + b;
            }
          } else {
#pragma unroll
            for (uint32_t b = WAVE / 2; b >= 1; b >>= 1) {
              const E v = lget_b<E>(WIN, ((h4 + 4u * (pos + b - 1u)) & m4w) | lw4);
              pos = (lt(y, v) || (int)(pos + b) > want) ? pos : pos + b;
            }
          }
          sk = (lane == ky || (full && !lt(y, lastv))) ? F : (int)pos;
        }
      } else if (live && want > 0) {
        if (lane == ky) {
          sk = F;
        } else {
          int l = 0, h = want;
          while (l < h) {
            const int mid = (l + h) >> 1;
            const E v = lget_b<E>(WIN, ((h4 + 4u * (uint32_t)mid) & m4w) | lw4);
            const bool tk = !lt(y, v);  // v <= y
            l = tk ? mid + 1 : l;
            h = tk ? h : mid;
          }
          sk = l;
        }
      }
      // gallop when no other list takes anything: list ky's run continues
      // past its window up to the smallest head z of the other lists
      // (elements <= z). Pairs also gallop when exactly one other list j
      // takes, and all it takes precedes list ky's head: j's elements go
      // out first, straight from its window, then ky gallops. That is a
      // list handing over to the next one where equal keys are ordered by
      // value across lists (each list's values in a range of their own, as
      // when values are input positions): the step would otherwise merge a
      // whole tile for a few outputs. (Keys without values have no such
      // order among equal keys: every list takes its copies at once.)
      const unsigned long long tk = __ballot(live && lane != ky && sk > 0);
      if (E::HAS_VAL && tk != 0 && (tk & (tk - 1)) == 0) {
        const int pj = __ffsll((long long)tk) - 1;
        const int skj = __builtin_amdgcn_readlane(sk, pj);
        if (skj <= WAVE && skj < __builtin_amdgcn_readlane(want, pj)) {
          // list ky's head and list pj's last element taken (every lane
          // reads the same two slots)
          const int hj = __builtin_amdgcn_readlane(head, pj);
          const E hk = lget<E>(win, WIN, ky * F + __builtin_amdgcn_readlane(head, ky));
          const E lj = lget<E>(win, WIN, pj * F + ((hj + skj - 1) & (F - 1)));
          if (!lt(hk, lj)) {
            pre = skj;
            if (lane < pre) {
              const E x = lget<E>(win, WIN, pj * F + ((hj + lane) & (F - 1)));
              const int64_t o = obase + lane;
              gstore<DT, CONV_OUT>(a.out_keys, a.out_vals, o, x, a.flip);
              if (a.samp_out_keys && ((uint32_t)o & (uint32_t)(SAMP - 1)) == 0)
                gstore<DT, false>(a.samp_out_keys, a.samp_out_vals, o >> SAMP_SHIFT, x);
            }
            if (lane == pj) {
              cur += pre;
              head = (head + pre) & (F - 1);
              want -= pre;
            }
            obase += pre;
          }
        }
      }
      gallop = tk == 0 || pre > 0;
    }

    if (gallop) {
      // Gallops, chained for pairs: list gk's run from its first element up
      // to the smallest head z of the other lists goes straight from input
      // to output; then list kz (the one whose head is z) goes on the same
      // way while its whole window precedes the next bound, without a step's
      // frontier and searches in between (runs of equal keys ordered by
      // value across lists pass from one list to the next like this).
      int gk = ky;
      for (bool first = true;; first = false) {
        E z;
        int kz;
        const bool ne = live && want > 0 && lane != gk;
        frontier_min<E>(ne ? lget<E>(win, WIN, lane * F + head) : Limits<E>::hi(), ne ? lane : WAVE, Kp, z, kz);
        if (!first && kz != WAVE) {
          // list gk goes on only if its window's last element is <= z
          const int hg = __builtin_amdgcn_readlane(head, gk), wg = __builtin_amdgcn_readlane(want, gk);
          if (lt(z, lget<E>(win, WIN, gk * F + ((hg + wg - 1) & (F - 1))))) break;
        }
        int gfill = -1;  // list gk's next window filled from registers (its count), or -1
        // list gk's run, from cur, that precedes z goes straight from input
        // to output: 4 elements per lane in flight, each batch checked
        // against z, up to the first element past z (the run is sorted, so
        // the check holds for a prefix) or the end of the list. One memory
        // round trip per 256 elements and none to find the end first.
        const int64_t src = uniform64(__shfl(cur, gk, WAVE));
        const int64_t lim = uniform64(__shfl(endp, gk, WAVE)) - src;
        int64_t pos = 0;
        for (;;) {
          E g[4];
          bool ok[4];
          const int rel = clamp_count_u(lim - pos, 4 * WAVE);  // elements left in the run, up to a batch
#pragma unroll
          for (int t = 0; t < 4; ++t) {
            const int64_t idx = pos + t * WAVE + lane;
            g[t] = t * WAVE + lane < rel ? gload<DT, false>(a.in_keys, a.in_vals, src + idx) : Limits<E>::hi();
          }
#pragma unroll
          for (int t = 0; t < 4; ++t) {
            const bool valid = t * WAVE + lane < rel;
            ok[t] = valid && (kz == WAVE || !lt(z, g[t]));  // g <= z
          }
          int f = 4 * WAVE;  // batch index (t*64 + lane order) of the first element not taken
#pragma unroll
          for (int t = 0; t < 4; ++t) {
            const unsigned long long nb = __ballot(!ok[t]);
            if (f == 4 * WAVE && nb != 0) f = t * WAVE + __ffsll((long long)nb) - 1;
          }
#pragma unroll
          for (int t = 0; t < 4; ++t) {
            const int64_t o = obase + pos + t * WAVE + lane;
            if (t * WAVE + lane < f) {
              gstore<DT, CONV_OUT>(a.out_keys, a.out_vals, o, g[t], a.flip);
              if (a.samp_out_keys && ((uint32_t)o & (uint32_t)(SAMP - 1)) == 0)
                gstore<DT, false>(a.samp_out_keys, a.samp_out_vals, o >> SAMP_SHIFT, g[t]);
            }
          }
          if (f < 4 * WAVE) {
            // the batch's elements past the run are list gk's next ones:
            // when they cover its next window, they go into it from
            // registers (slot i = element i, head 0) and the window needs
            // no reload from memory
            const int64_t rest = lim - pos - f;  // list gk's elements after the run
            const int need = rest < (int64_t)F ? (int)rest : F;
            if (F <= 4 * WAVE && rel - f >= need) {
#pragma unroll
              for (int t = 0; t < 4; ++t) {
                const int i = t * WAVE + lane - f;
                if (i >= 0 && i < need) lput<E>(win, WIN, gk * F + i, g[t]);
              }
              gfill = need;
              wave_lds_sync();
            }
          }
          pos += f;
          if (f < 4 * WAVE) break;
        }
        const int64_t glen = pos;
        // list gk restarts with a window from its new position (filled
        // above from the gallop's registers, or reloaded); the others keep
        // theirs (list pj of a hand-over reloads the slots it gave up)
        int c2 = want;
        if (lane == gk) {
          cur += glen;
          head = 0;
          c2 = gfill < 0 ? 0 : gfill;
        }
        obase += glen;
        const int w2 = live ? clamp_count(endp - cur, F) : 0;
        want = w2;
        if (gfill < 0 || pre > 0) {
          refill(cur, head, c2, w2);
          wait_dma();
          pre = 0;
        }
        // no other list holds anything (list gk ran to its end); keys
        // without values do not chain (their equal keys need no hand-over
        // from list to list, and their step stays as compiled without it)
        if (!E::HAS_VAL || kz == WAVE) break;
        gk = kz;
      }
      continue;
    }

    // ---- elements taken this step, and the state after it
    // S: row sums by DPP butterflies, then the four rows' totals (no LDS)
    int S;
    {
      uint32_t t = (uint32_t)sk;
      t += dpp_mov<0xB1>(t);   // quad_perm [1,0,3,2]
      t += dpp_mov<0x4E>(t);   // quad_perm [2,3,0,1]
      t += dpp_mov<0x141>(t);  // row_half_mirror
      t += dpp_mov<0x140>(t);  // row_mirror
      S = __builtin_amdgcn_readlane((int)t, 0) + __builtin_amdgcn_readlane((int)t, 16) +
          __builtin_amdgcn_readlane((int)t, 32) + __builtin_amdgcn_readlane((int)t, 48);
    }
    const int64_t cur2 = cur + sk;
    const int head2 = (head + sk) & (F - 1);
    const int cnt2 = want - sk;
    const int want2 = live ? clamp_count(endp - cur2, F) : 0;
    // outputs (o < S) in the store layout above
    E ov[NREG];
    if (Kp == 1) {
      // one run: it is window 0 (list 0's, held by lane 0), in order from its head
      const int h0 = __builtin_amdgcn_readlane(head, 0);
#pragma unroll
      for (int q = 0; q < NREG; ++q) {
        const int o = VEC ? (q >> 2) * 256 + 4 * lane + (q & 3) : q * WAVE + lane;
        ov[q] = lget<E>(win, WIN, (h0 + o) & (F - 1));
      }
      __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0): window reads done before the DMA reuses the slots
      refill(cur2, head2, cnt2, want2);
    } else {
      // register tile: lane l holds slots l*NREG .. l*NREG+NREG-1, all in run r
      const int r = tr;
      const int i0 = (lane * NREG) & (F - 1);
      E v[NREG];
      // slot q: logical position hd + i0 + q of run r's window, which starts
      // at word r*F (a multiple of F, so the wrapped offset ORs into it)
      const uint32_t rb = lds_addr(win) + (uint32_t)(r * F) * 4u, m4 = (uint32_t)F * 4u - 1u;
      uint32_t t4 = (uint32_t)(hd + i0) * 4u;
      asm("" : "+v"(t4));  // keeps the per-slot address at one add and one and-or
#pragma unroll
      for (int q = 0; q < NREG; ++q) v[q] = lget_b<E>(WIN, ((t4 + 4u * (uint32_t)q) & m4) | rb);
      // The untaken elements of a window are > y, so they sort behind the S
      // taken ones and may stay; only slots past a list's window contents
      // (a list near its end) need the filler.
      if (__ballot(i0 + NREG > wr) != 0) {
#pragma unroll
        for (int q = 0; q < NREG; ++q) v[q] = sel(i0 + q < wr, v[q], Limits<E>::hi());
      }
      __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0): window reads done before the DMA reuses the slots
      refill(cur2, head2, cnt2, want2);
      // merge the Kp sorted blocks of F slots: levels from block F up to WIN
      const int b0 = __builtin_ctz((unsigned)(F / NREG));
      switch (b0) {
        case 0: merge_level<0, XC_NOLDS>(v, lane); [[fallthrough]];
        case 1: merge_level<1, XC_NOLDS>(v, lane); [[fallthrough]];
        case 2: merge_level<2, XC_NOLDS>(v, lane); [[fallthrough]];
        case 3: merge_level<3, XC_NOLDS>(v, lane); [[fallthrough]];
        case 4: merge_level<4, XC_NOLDS>(v, lane); [[fallthrough]];
        default: merge_level<5, XC_NOLDS>(v, lane);
      }
      if constexpr (VEC) {
        // transpose through LDS in 16-byte chunks so that every lane stores
        // four consecutive outputs: lane l's slots go to row l (NREG words),
        // chunk k at chunk position k ^ xswz(l). The swizzle puts the 8 (or
        // 16) lanes that one LDS cycle serves on distinct banks, for the
        // writes and for the reads of output rows of 256.
#pragma unroll
        for (int k = 0; k < XC; ++k) {
          const u32x4 c = {word(v[4 * k], 0), word(v[4 * k + 1], 0), word(v[4 * k + 2], 0), word(v[4 * k + 3], 0)};
          *(u32x4*)(xp + lane * NREG + 4 * (k ^ xswz<XC>(lane))) = c;
        }
        wave_lds_sync();
        // output o = 256u + 4*lane + t sits in tile lane L = o / NREG, chunk
        // (o % NREG) / 4; row u adds 256 words and leaves the swizzle alone
        const int L = (4 * lane) / NREG;
        const int rb = L * NREG + 4 * ((lane % XC) ^ xswz<XC>(L));
#pragma unroll
        for (int u = 0; u < XC; ++u) {
          const u32x4 c = *(const u32x4*)(xp + u * 256 + rb);
#pragma unroll
          for (int t = 0; t < 4; ++t) {
            const uint32_t wd[1] = {c[t]};
            ov[4 * u + t] = from_words<E>(wd);
          }
        }
      } else {
        // transpose through LDS (rows padded to NREG + 1 words against bank
        // conflicts) so that the one-element stores are coalesced; pairs in
        // two halves of the tile: output o = u*64 + lane sits in tile lane
        // o / NREG, register o % NREG, so outputs u < NREG/2 come from tile
        // lanes 0-31 and the others from 32-63 (row = tile lane - 32h)
        constexpr int UH = NREG / XH;
        constexpr int RH = WAVE / XH;  // tile lanes per part
        const int xb = (lane / NREG) * XS + (lane % NREG);
#pragma unroll
        for (int h = 0; h < XH; ++h) {
          if (XH == 1 || lane / RH == h) {
#pragma unroll
            for (int w = 0; w < W; ++w) {
#pragma unroll
              for (int q = 0; q < NREG; ++q) xp[w * XW + (lane % RH) * XS + q] = word(v[q], w);
            }
          }
          wave_lds_sync();
#pragma unroll
          for (int u = 0; u < UH; ++u) {
            const int slot = xb + u * (WAVE / NREG) * XS;
            uint32_t wd[W];
#pragma unroll
            for (int w = 0; w < W; ++w) wd[w] = xp[w * XW + slot];
            ov[h * UH + u] = from_words<E>(wd);
          }
          if (XH > 1) wave_lds_sync();  // this part's reads before the next part's writes
        }
      }
    }
    wait_dma();
    // the next round's partition samples (every SAMP-th output, internal
    // bits) of this step's outputs, from registers
    if (a.samp_out_keys) {
      if constexpr (VEC) {
        // output obase + o with o = 256u + 4*lane + t: the multiples of SAMP
        // (= 256) are o = s0 + 256u, all in lane s0/4, register 4u + s0%4
        const int s0 = (int)((uint64_t)(-obase) & (uint64_t)(SAMP - 1));
        const int ns = S > s0 ? (S - s0 + SAMP - 1) >> SAMP_SHIFT : 0;  // samples in this step (<= XC)
        if (ns > 0 && lane == (s0 >> 2)) {
          const int64_t si = (obase + s0) >> SAMP_SHIFT;  // index of the first
          auto put = [&](auto tc) {
            constexpr int T = decltype(tc)::value;
            gstore<DT, false>(a.samp_out_keys, a.samp_out_vals, si, ov[T]);
#pragma unroll
            for (int u = 1; u < XC; ++u)
              if (u < ns) gstore<DT, false>(a.samp_out_keys, a.samp_out_vals, si + u, ov[4 * u + T]);
          };
          switch (s0 & 3) {  // wave-uniform
            case 0: put(std::integral_constant<int, 0>{}); break;
            case 1: put(std::integral_constant<int, 1>{}); break;
            case 2: put(std::integral_constant<int, 2>{}); break;
            default: put(std::integral_constant<int, 3>{}); break;
          }
        }
      } else {
#pragma unroll
        for (int u = 0; u < NREG; ++u) {
          const int o = u * WAVE + lane;
          if (o < S && (((uint32_t)obase + (uint32_t)o) & (uint32_t)(SAMP - 1)) == 0)
            gstore<DT, false>(a.samp_out_keys, a.samp_out_vals, (obase + o) >> SAMP_SHIFT, ov[u]);
        }
      }
    }
    if constexpr (VEC) {
      // rows below S/256 are full (a wave-uniform test); in the partial row a
      // lane stores its four outputs at once, or one by one at the end
      const int nfull = S >> 8;
#pragma unroll
      for (int u = 0; u < XC; ++u) {
        const int o = u * 256 + 4 * lane;
        if (u < nfull) {
          gstore4<DT, CONV_OUT>(a.out_keys, a.out_vals, obase, o, &ov[4 * u], a.flip);
        } else if (u == nfull) {
          if (o + 4 <= S) {
            gstore4<DT, CONV_OUT>(a.out_keys, a.out_vals, obase, o, &ov[4 * u], a.flip);
          } else if (o < S) {
            // the lane holding the last 1-3 outputs: an 8-byte store of the
            // first two, then the remaining one
            const int rem = S - o;
            if (rem & 2) gstore2<DT, CONV_OUT>(a.out_keys, obase, o, &ov[4 * u], a.flip);
            if (rem & 1) {
              // a lane-mask select: a plain ?: over two ov[] entries becomes
              // an indexed access that moves ov into scratch memory
              E e;
              e.k = mask_sel(__ballot(rem & 2), word(ov[4 * u + 2], 0), word(ov[4 * u], 0));
              gstore<DT, CONV_OUT>(a.out_keys, a.out_vals, obase + o + (rem & 2), e, a.flip);
            }
          }
        }
      }
    } else {
      const int nu = (S + WAVE - 1) / WAVE;  // store rows, the last one partial
#pragma unroll
      for (int u = 0; u < NREG; ++u) {
        const int o = u * WAVE + lane;
        if (u < nu - 1 || (u == nu - 1 && o < S))
          gstore<DT, CONV_OUT>(a.out_keys, a.out_vals, obase + o, ov[u], a.flip);
      }
    }

    // ---- advance
    obase += S;
    cur = cur2;
    head = head2;
    want = want2;
  }
  // done: ranks below every merger still running
  if constexpr (PEERS) *(volatile __attribute__((address_space(3))) uint32_t*)(size_t)wp.mine = 0u;
}

// Plain sort: one merger per wavefront over the uniform task grid.
// PEERS (the round's mergers all resident at once): a workgroup holds a
// CU's resident mergers (blockDim.x / 64 of them), which rank their
// progress against the others on their SIMD. Otherwise (more mergers than
// resident slots: later ones start as earlier ones finish) workgroups of
// WM_WAVES mergers, so that a finished workgroup frees its slots early,
// with the progress bands.
template <int DT, bool CONV_OUT, int WIN, bool PEERS>
__global__ __launch_bounds__(WM_PLAIN_MAX* WAVE, 4) void k_wmerge(RoundArgs a) {
  using E = typename DType<DT>::E;
  extern __shared__ uint32_t lds[];
  const int lane = threadIdx.x & (WAVE - 1);
  const int nw = (int)(blockDim.x / WAVE);
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);  // wave-uniform
  const int64_t merger = (int64_t)blockIdx.x * nw + wv;
  uint32_t* win = wmerge_windows(lds, wv, E::W, WIN);
  uint32_t* xp = wmerge_transpose(lds, nw, wv, E::W, WIN);
  if constexpr (!PEERS) {
    if (merger >= a.tasks * a.m) return;  // whole wavefront leaves; no barriers here
    wmerge_one<DT, CONV_OUT, WIN, false>(a, plain_merger(a, merger), lane, win, xp, WavePeers{0u, 0u});
    return;
  }
  // every merger's SIMD (HW_ID bits 5:4); each merger's progress word is
  // the k-th of its SIMD's four, k = the mergers of lower index on that SIMD
  uint32_t* tab = lds + (size_t)nw * wmerge_lds_words_per_wave(E::W, WIN);
  const int simd = (int)((__builtin_amdgcn_s_getreg(4 | (0 << 6) | (31 << 11)) >> 4) & 3u);
  if (lane == 0) tab[wv] = (uint32_t)simd;
  if (wv == 0 && lane < 16) tab[nw + lane] = 0xFFFFFFFFu;
  __syncthreads();  // the workgroup's only barrier
  int k = 0;
  for (int w = 0; w < wv; ++w) k += (int)tab[w] == simd ? 1 : 0;
  k = __builtin_amdgcn_readfirstlane(k);  // wave-uniform (every lane read the same words)
  if (merger >= a.tasks * a.m) return;  // whole wavefront leaves (its progress word stays all ones)
  const uint32_t grp = lds_addr(tab + nw + 4 * simd);
  // (k < 4 at 4 wavefronts per SIMD; were it not, two mergers would share a
  // word, which costs only priority accuracy)
  wmerge_one<DT, CONV_OUT, WIN, true>(a, plain_merger(a, merger), lane, win, xp, WavePeers{grp, grp + 4u * (k & 3)});
}

// Batched segments: a persistent launch; every wavefront takes merger
// descriptors from a.items through the a.queue counter until the list is
// exhausted (every wavefront reaches that exit). The counter starts at zero
// (cleared by the launch before).
template <int DT, bool CONV_OUT, int WIN>
__global__ __launch_bounds__(WM_WAVES* WAVE, 4) void k_wmerge_persistent(RoundArgs a) {
  using E = typename DType<DT>::E;
  extern __shared__ uint32_t lds[];
  const int lane = threadIdx.x & (WAVE - 1);
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);  // wave-uniform
  uint32_t* win = wmerge_windows(lds, wv, E::W, WIN);
  uint32_t* xp = wmerge_transpose(lds, WM_WAVES, wv, E::W, WIN);
  const WavePeers wp{0u, 0u};
  for (;;) {
    unsigned int it = 0;
    if (lane == 0) it = atomicAdd(a.queue, 1u);
    // wave-uniform in a scalar register, so that the work item's fields
    // load with scalar loads and stay out of the vector registers
    it = (unsigned int)__builtin_amdgcn_readfirstlane((int)it);
    if ((int64_t)it >= a.n_items) break;
    wmerge_one<DT, CONV_OUT, WIN, false>(a, load_merger(a.items + it), lane, win, xp, wp);
    wave_lds_sync();
  }
}

}  // namespace gmms
