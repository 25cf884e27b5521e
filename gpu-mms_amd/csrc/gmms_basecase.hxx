// Base case: every 1024-element tile sorted by one 64-lane wavefront, in
// registers. Replaces the reference's squareSort (src/basecase/
// squareSort.hxx:62-131), which sorts a 32x32 tile with one 32-lane warp via
// bitonicSort32 + merge32..merge512 (src/basecase/sortRowMajor.hxx:72-220)
// and selection-sorts the N%1024 tail on one thread (:69-88).
//
// Layout: element e of the tile lives in lane e/16, register e%16 (16 keys per
// lane). The network is bitonic sort in "flip" form (every compare-exchange
// puts the minimum at the lower index):
//   * 16 registers per lane: Batcher odd-even merge sort (63 exchanges),
//   * for block sizes 32..1024: one mirror stage across lanes (lane ^ (2^b-1),
//     register r <-> 15-r), half-cleaners across lanes at lane distance
//     2^(b-1)..1 (DPP / ds_swizzle / ds_bpermute), then in-register
//     half-cleaners at distances 8,4,2,1.
// The ragged last tile is padded with the maximum element in registers, not
// in memory (the reference's pad.hxx writes MAXVAL-1 into the array and its
// sentinel collides with real keys; here a padded slot is simply never stored).
// Loads are fully coalesced (lane + 64*r); the sorted tile goes back to
// memory through a bank-padded LDS transpose so the stores are coalesced too.
#pragma once
#include <type_traits>
#include "gmms_elem.hxx"

namespace gmms {

constexpr int TILE = 1024;
constexpr int TPL = TILE / WAVE;  // keys per lane

// Batcher odd-even merge sort on N registers (iterative form, Knuth 5.3.4).
template <class E, int N>
__device__ __forceinline__ void oddeven_sort(E (&v)[N]) {
#pragma unroll
  for (int p = 1; p < N; p <<= 1) {
#pragma unroll
    for (int k = p; k >= 1; k >>= 1) {
#pragma unroll
      for (int j = k % p; j + k < N; j += 2 * k) {
#pragma unroll
        for (int i = 0; i < k; ++i) {
          if (i + j + k < N && (i + j) / (2 * p) == (i + j + k) / (2 * p)) cswap(v[i + j], v[i + j + k]);
        }
      }
    }
  }
}

// in-register half-cleaners, distances N/2 .. 1
template <class E, int N>
__device__ __forceinline__ void clean_regs(E (&v)[N]) {
#pragma unroll
  for (int d = N / 2; d >= 1; d >>= 1) {
#pragma unroll
    for (int r = 0; r < N; ++r) {
      if ((r & d) == 0) cswap(v[r], v[r + d]);
    }
  }
}

// mirror stage of block size 16 * 2^(B+1): partner lane ^ (2^(B+1)-1), reg N-1-r
template <int B, class E, int N>
__device__ __forceinline__ void flip_lanes(E (&v)[N], int lane) {
  constexpr int M = (2 << B) - 1;
  const bool low = ((lane >> B) & 1) == 0;
  const uint32_t hsel = low ? 0u : 0xFFFFFFFFu;
  E p[N];
#pragma unroll
  for (int r = 0; r < N; ++r) p[r] = xlane<M>(v[N - 1 - r]);
#pragma unroll
  for (int r = 0; r < N; ++r) v[r] = lane_ce(v[r], p[r], low, hsel);
}

// 32-bit keys, half-cleaner at lane distance 4, four registers: DPP has no
// xor-4 pattern, but a lane pair at distance 4 lies in one row, the lower
// lane in bank 0 or 2 (lanes 0-3, 8-11 of the row), the upper in bank 1 or 3.
// The lower lanes take min(x, lane + 4) by row_ror:12, the upper lanes
// max(x, lane - 4) by row_ror:4, each instruction writing only its banks:
// two VALU instructions per register and no LDS round trip. The s_nop
// before covers the DPP read of registers written just before the block,
// the one after a DPP read of its last result just after it.
__device__ __forceinline__ void clean4_u32x4(uint32_t& a, uint32_t& b, uint32_t& c, uint32_t& d) {
  uint32_t ta, tb, tc, td;
  asm("s_nop 1\n\t"
      "v_min_u32_dpp %0, %4, %4 row_ror:12 row_mask:0xf bank_mask:0x5\n\t"
      "v_min_u32_dpp %1, %5, %5 row_ror:12 row_mask:0xf bank_mask:0x5\n\t"
      "v_min_u32_dpp %2, %6, %6 row_ror:12 row_mask:0xf bank_mask:0x5\n\t"
      "v_min_u32_dpp %3, %7, %7 row_ror:12 row_mask:0xf bank_mask:0x5\n\t"
      "v_max_u32_dpp %0, %4, %4 row_ror:4 row_mask:0xf bank_mask:0xa\n\t"
      "v_max_u32_dpp %1, %5, %5 row_ror:4 row_mask:0xf bank_mask:0xa\n\t"
      "v_max_u32_dpp %2, %6, %6 row_ror:4 row_mask:0xf bank_mask:0xa\n\t"
      "v_max_u32_dpp %3, %7, %7 row_ror:4 row_mask:0xf bank_mask:0xa\n\t"
      "s_nop 1"
      : "=&v"(ta), "=&v"(tb), "=&v"(tc), "=&v"(td)
      : "v"(a), "v"(b), "v"(c), "v"(d));
  a = ta;
  b = tb;
  c = tc;
  d = td;
}

// 32-bit keys, half-cleaner at lane distance 16 or 32 on registers a, b:
// v_permlane16_swap (32_swap) exchanges the odd rows of 16 (the upper 32
// lanes) of a with the even rows (the lower 32 lanes) of b, so that every
// lower lane's element and its upper partner meet in one lane, the lower one
// in a; min and max there, and the same swap puts them back. Four VALU
// instructions per two registers and no LDS round trip.
template <int D>
__device__ __forceinline__ void clean_rows_u32x2(uint32_t& a, uint32_t& b) {
  static_assert(D == 16 || D == 32, "row swaps");
  const auto s = D == 16 ? __builtin_amdgcn_permlane16_swap(a, b, false, false)
                         : __builtin_amdgcn_permlane32_swap(a, b, false, false);
  const uint32_t lo = s[0] < s[1] ? s[0] : s[1], hi = s[0] < s[1] ? s[1] : s[0];
  const auto t = D == 16 ? __builtin_amdgcn_permlane16_swap(lo, hi, false, false)
                         : __builtin_amdgcn_permlane32_swap(lo, hi, false, false);
  a = t[0];
  b = t[1];
}

// How the cross-lane stages of a network exchange values:
//   XC_LDS   DPP where a pattern exists, ds_swizzle / ds_bpermute otherwise:
//            the fewest VALU instructions, for kernels bound by VALU issue
//            (the base case);
//   XC_NOLDS 32-bit keys use DPP, bank-masked DPP and row swaps only, one or
//            two VALU instructions more per register and stage but no LDS
//            round trip in the dependency chain, for kernels bound by the
//            latency of that chain (the merge step).
enum : int { XC_LDS = 0, XC_NOLDS = 1 };

// half-cleaner across lanes at lane distance D
template <int D, int X = XC_LDS, class E, int N>
__device__ __forceinline__ void clean_lanes(E (&v)[N], int lane) {
  if constexpr (X == XC_NOLDS && D == 4 && std::is_same<E, E32>::value && N % 4 == 0) {
#pragma unroll
    for (int r = 0; r < N; r += 4) clean4_u32x4(v[r].k, v[r + 1].k, v[r + 2].k, v[r + 3].k);
    return;
  }
  if constexpr (X == XC_NOLDS && (D == 16 || D == 32) && std::is_same<E, E32>::value && N % 2 == 0) {
#pragma unroll
    for (int r = 0; r < N; r += 2) clean_rows_u32x2<D>(v[r].k, v[r + 1].k);
    return;
  }
  const bool low = (lane & D) == 0;
  const uint32_t hsel = low ? 0u : 0xFFFFFFFFu;
#pragma unroll
  for (int r = 0; r < N; ++r) {
    E p = xlane<D>(v[r]);
    v[r] = lane_ce(v[r], p, low, hsel);
  }
}

template <int B, int X = XC_LDS, class E, int N>
__device__ __forceinline__ void lane_cleaners(E (&v)[N], int lane) {
  if constexpr (B >= 6) clean_lanes<32, X>(v, lane);
  if constexpr (B >= 5) clean_lanes<16, X>(v, lane);
  if constexpr (B >= 4) clean_lanes<8, X>(v, lane);
  if constexpr (B >= 3) clean_lanes<4, X>(v, lane);
  if constexpr (B >= 2) clean_lanes<2, X>(v, lane);
  if constexpr (B >= 1) clean_lanes<1, X>(v, lane);
}

template <int B, int X = XC_LDS, class E, int N>
__device__ __forceinline__ void merge_level(E (&v)[N], int lane) {
  flip_lanes<B>(v, lane);
  lane_cleaners<B, X>(v, lane);
  clean_regs(v);
}

// Sort 1024 elements held as v[16] per lane; result: lane l holds ranks
// 16l .. 16l+15 in ascending register order.
template <class E>
__device__ __forceinline__ void sort_tile(E (&v)[TPL], int lane) {
  oddeven_sort(v);
  merge_level<0>(v, lane);
  merge_level<1>(v, lane);
  merge_level<2>(v, lane);
  merge_level<3>(v, lane);
  merge_level<4>(v, lane);
  merge_level<5>(v, lane);
}

constexpr int BASE_WAVES = 4;                   // waves (tiles) per workgroup
constexpr int XPOSE_STRIDE = TPL + 1;            // bank padding
constexpr int XPOSE_WORDS = WAVE * XPOSE_STRIDE; // per wave per word plane

// Exchange with another wavefront of the workgroup through LDS: lane l,
// register r receives the partner wavefront's element at lane 63-l, register
// 15-r (MIRROR: the flip stage across two 1024-blocks) or at lane l,
// register r (a half-cleaner at distance 1024), and keeps the minimum if LOW.
// `buf` holds BASE_WAVES * W * XPOSE_WORDS words; every wavefront of the
// workgroup calls this together (two workgroup barriers).
template <bool MIRROR, class E>
__device__ __forceinline__ void wave_exchange(E (&v)[TPL], uint32_t* buf, int wave, int partner, bool low, int lane) {
  constexpr int W = E::W;
  __syncthreads();  // the previous use of buf is over
  uint32_t* mine = buf + wave * W * XPOSE_WORDS;
  if constexpr (W == 1) {
    // 32-bit keys: rows of 16 words in swizzled 16-byte chunks (chunk k of
    // row l at chunk position k ^ xswz(l)), read back 16 bytes at a time
#pragma unroll
    for (int k = 0; k < TPL / 4; ++k) {
      const u32x4 c = {v[4 * k].k, v[4 * k + 1].k, v[4 * k + 2].k, v[4 * k + 3].k};
      *(u32x4*)(mine + lane * TPL + 4 * (k ^ xswz<4>(lane))) = c;
    }
    __syncthreads();
    const uint32_t* theirs = buf + partner * W * XPOSE_WORDS;
    const int pl = MIRROR ? WAVE - 1 - lane : lane;
    const uint32_t hsel = low ? 0u : 0xFFFFFFFFu;
#pragma unroll
    for (int k = 0; k < TPL / 4; ++k) {
      // MIRROR: register 4k+t meets the partner's register 15-4k-t, the
      // reversed chunk 3-k
      const int pk = MIRROR ? TPL / 4 - 1 - k : k;
      const u32x4 c = *(const u32x4*)(theirs + pl * TPL + 4 * (pk ^ xswz<4>(pl)));
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        E p;
        p.k = c[MIRROR ? 3 - t : t];
        v[4 * k + t] = lane_ce(v[4 * k + t], p, low, hsel);
      }
    }
    return;
  }
#pragma unroll
  for (int w = 0; w < W; ++w) {
#pragma unroll
    for (int r = 0; r < TPL; ++r) mine[w * XPOSE_WORDS + lane * XPOSE_STRIDE + r] = word(v[r], w);
  }
  __syncthreads();
  const uint32_t* theirs = buf + partner * W * XPOSE_WORDS;
  const int pl = MIRROR ? WAVE - 1 - lane : lane;
  const uint32_t hsel = low ? 0u : 0xFFFFFFFFu;
#pragma unroll
  for (int r = 0; r < TPL; ++r) {
    const int pr = MIRROR ? TPL - 1 - r : r;
    uint32_t wd[W];
#pragma unroll
    for (int w = 0; w < W; ++w) wd[w] = theirs[w * XPOSE_WORDS + pl * XPOSE_STRIDE + pr];
    v[r] = lane_ce(v[r], from_words<E>(wd), low, hsel);
  }
}

// Tile load: elements base .. base + cnt - 1 into v (the maximum element
// past cnt). A full tile loads 16 bytes of keys per lane and load: for
// 32-bit keys registers 4k..4k+3 of lane l are elements 256k + 4l .. + 3
// (any placement works: the tile is sorted as a whole).
template <int DT, bool CONV>
__device__ __forceinline__ void tile_load(const void* keys, const uint32_t* vals, int64_t base, int cnt,
                                          uint32_t flip, typename DType<DT>::E (&v)[TPL], int lane) {
  using E = typename DType<DT>::E;
  if constexpr (E::W == 1) {
    if (cnt == TILE) {
#pragma unroll
      for (int k = 0; k < TPL / 4; ++k) gload4<DT, CONV>(keys, base, 256 * k + 4 * lane, &v[4 * k], flip);
      return;
    }
  } else {
    // 64-bit keys: two elements per 16-byte load, registers 2k, 2k+1 of
    // lane l are elements 128k + 2l, 128k + 2l + 1
    if (cnt == TILE) {
#pragma unroll
      for (int k = 0; k < TPL / 2; ++k) gload2<DT, CONV>(keys, vals, base, 128 * k + 2 * lane, &v[2 * k], flip);
      return;
    }
  }
#pragma unroll
  for (int r = 0; r < TPL; ++r) {
    const int idx = r * WAVE + lane;
    v[r] = idx < cnt ? gload<DT, CONV>(keys, vals, base + idx, flip) : Limits<E>::hi();
  }
}

// Tile store of the sorted tile (lane l holds ranks 16l .. 16l+15) to
// base .. base + cnt - 1 through the wavefront's LDS region `my`
// (E::W * XPOSE_WORDS words), so that the stores are coalesced.
template <int DT, bool CONV>
__device__ __forceinline__ void tile_store(void* keys, uint32_t* vals, int64_t base, int cnt, uint32_t flip,
                                           const typename DType<DT>::E (&v)[TPL], uint32_t* my, int lane) {
  using E = typename DType<DT>::E;
  if constexpr (E::W == 1) {
    // 16-byte chunks, chunk k of lane l at chunk position k ^ xswz(l) of its
    // 16-word row (no bank conflicts); lane l then reads and stores the four
    // consecutive ranks 256u + 4l .. + 3 of each row u
#pragma unroll
    for (int k = 0; k < TPL / 4; ++k) {
      const u32x4 c = {v[4 * k].k, v[4 * k + 1].k, v[4 * k + 2].k, v[4 * k + 3].k};
      *(u32x4*)(my + lane * TPL + 4 * (k ^ xswz<4>(lane))) = c;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int L = lane / 4;  // the lane whose row holds ranks 4*lane .. 4*lane + 3
    const int rb = L * TPL + 4 * ((lane & 3) ^ xswz<4>(L));
#pragma unroll
    for (int u = 0; u < TPL / 4; ++u) {
      const u32x4 c = *(const u32x4*)(my + u * 256 + rb);
      E e4[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) e4[t].k = c[t];
      const int o = u * 256 + 4 * lane;
      if (cnt == TILE) {
        gstore4<DT, CONV>(keys, vals, base, o, e4, flip);
      } else {
#pragma unroll
        for (int t = 0; t < 4; ++t)
          if (o + t < cnt) gstore<DT, CONV>(keys, vals, base + o + t, e4[t], flip);
      }
    }
  } else {
#pragma unroll
    for (int w = 0; w < E::W; ++w) {
#pragma unroll
      for (int r = 0; r < TPL; ++r) my[w * XPOSE_WORDS + lane * XPOSE_STRIDE + r] = word(v[r], w);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (cnt == TILE) {
      // lane l stores the four consecutive ranks 256u + 4l .. + 3 of each
      // row u: two 16-byte key stores (and one 16-byte value store)
#pragma unroll
      for (int u = 0; u < TPL / 4; ++u) {
        const int o = u * 256 + 4 * lane;
        E e4[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const int slot = (o / TPL) * XPOSE_STRIDE + (o % TPL) + t;  // o % TPL + t < TPL
          uint32_t w[E::W];
#pragma unroll
          for (int q = 0; q < E::W; ++q) w[q] = my[q * XPOSE_WORDS + slot];
          e4[t] = from_words<E>(w);
        }
        gstore4<DT, CONV>(keys, vals, base, o, e4, flip);
      }
      return;
    }
#pragma unroll
    for (int r = 0; r < TPL; ++r) {
      const int e = r * WAVE + lane;  // rank
      const int slot = (e / TPL) * XPOSE_STRIDE + (e % TPL);
      uint32_t w[E::W];
#pragma unroll
      for (int q = 0; q < E::W; ++q) w[q] = my[q * XPOSE_WORDS + slot];
      if (e < cnt) gstore<DT, CONV>(keys, vals, base + e, from_words<E>(w), flip);
    }
  }
}

// In-place base case over keys[0..n): CONV_IN maps raw -> internal on load,
// CONV_OUT maps internal -> raw on store. RUN4: the workgroup's four sorted
// tiles are merged into one sorted run of 4096 (two more merge levels; the
// flip and distance-1024 stages go through LDS between wavefronts, the rest
// are the in-wavefront stages of merge_level).
template <int DT, bool CONV_IN, bool CONV_OUT, bool RUN4 = false>
__global__ __launch_bounds__(BASE_WAVES * WAVE) void k_basecase(void* keys, uint32_t* vals, int64_t n,
                                                               uint32_t flip) {
  using E = typename DType<DT>::E;
  __shared__ __align__(16) uint32_t xp[BASE_WAVES * E::W * XPOSE_WORDS];
  const int lane = threadIdx.x & (WAVE - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
  const int64_t tile = (int64_t)blockIdx.x * BASE_WAVES + wave;
  const int64_t base = tile * TILE;
  // without RUN4 there is no workgroup barrier below and an idle wave leaves;
  // with it, a wave past the end sorts a tile of maxima and stores nothing
  if (!RUN4 && base >= n) return;
  const int64_t rest = n - base;
  const int cnt = rest <= 0 ? 0 : (int)(rest < TILE ? rest : TILE);

  E v[TPL];
  tile_load<DT, CONV_IN>(keys, vals, base, cnt, flip, v, lane);

  sort_tile(v, lane);

  if constexpr (RUN4) {
    static_assert(BASE_WAVES == 4, "a run of 4096 is four tiles");
    // 1024 -> 2048: flip across wavefront pairs, then the in-wavefront cleaners
    wave_exchange<true>(v, xp, wave, wave ^ 1, (wave & 1) == 0, lane);
    lane_cleaners<6>(v, lane);
    clean_regs(v);
    // 2048 -> 4096: flip across the two halves, the cleaner at distance 1024,
    // then the in-wavefront cleaners
    wave_exchange<true>(v, xp, wave, 3 - wave, wave < 2, lane);
    wave_exchange<false>(v, xp, wave, wave ^ 1, (wave & 1) == 0, lane);
    lane_cleaners<6>(v, lane);
    clean_regs(v);
    __syncthreads();  // xp is reused by the store transpose below
  }

  tile_store<DT, CONV_OUT>(keys, vals, base, cnt, flip, v, xp + wave * E::W * XPOSE_WORDS, lane);
}

// ---------------------------------------------------------------------
// Runs of 4096 32-bit keys sorted by one wavefront alone: 64 keys per lane
// (element e in lane e/64, register e%64). The network is the same bitonic
// sort in flip form, but with 6 of the 12 index bits in registers 57 of its
// 78 stages are in-register compare-exchanges (one VALU instruction per
// element) and 21 cross lanes, against 42 and 36 with 16 keys per lane and
// four wavefronts per run, and no workgroup barrier or LDS exchange is
// left. The in-lane sort of 64 keys is Batcher's odd-even merge sort (543
// exchanges, against 672 for the bitonic stages it replaces). About 1360
// VALU instructions per 1024 keys, against 1665.
constexpr int RUN = 4096;
constexpr int RPL = RUN / WAVE;  // keys per lane

// mirror stage across lanes for many registers: partners fetched two at a
// time (registers r and N-1-r exchange with the mirrored lane's N-1-r and
// r), so no array of N partners is live
template <int B, class E, int N>
__device__ __forceinline__ void flip_lanes_pairs(E (&v)[N], int lane) {
  constexpr int M = (2 << B) - 1;
  const bool low = ((lane >> B) & 1) == 0;
  const uint32_t hsel = low ? 0u : 0xFFFFFFFFu;
#pragma unroll
  for (int r = 0; r < N / 2; ++r) {
    const E a = xlane<M>(v[N - 1 - r]);
    const E b = xlane<M>(v[r]);
    v[r] = lane_ce(v[r], a, low, hsel);
    v[N - 1 - r] = lane_ce(v[N - 1 - r], b, low, hsel);
  }
}

// One run of cnt (1..4096) 32-bit keys at src + base, sorted into
// dst + base by the calling wavefront.
template <int DT, bool CONV_IN, bool CONV_OUT>
__device__ __forceinline__ void sort_run(const void* src, void* dst, int64_t base, int cnt, uint32_t flip,
                                         int lane) {
  using E = typename DType<DT>::E;
  static_assert(E::W == 1, "32-bit keys");
  E v[RPL];
  if (cnt == RUN) {
    // registers 4k..4k+3 of lane l: elements 256k + 4l .. + 3 (16-byte loads;
    // any placement works, the run is sorted as a whole)
#pragma unroll
    for (int k = 0; k < RPL / 4; ++k) gload4<DT, CONV_IN>(src, base, 256 * k + 4 * lane, &v[4 * k], flip);
  } else {
#pragma unroll
    for (int r = 0; r < RPL; ++r) {
      const int idx = r * WAVE + lane;
      v[r] = idx < cnt ? gload<DT, CONV_IN>(src, nullptr, base + idx, flip) : Limits<E>::hi();
    }
  }

  oddeven_sort(v);
  auto level = [&](auto bc) {
    constexpr int B = decltype(bc)::value;
    flip_lanes_pairs<B>(v, lane);
    lane_cleaners<B>(v, lane);
    clean_regs(v);
  };
  level(std::integral_constant<int, 0>{});
  level(std::integral_constant<int, 1>{});
  level(std::integral_constant<int, 2>{});
  level(std::integral_constant<int, 3>{});
  level(std::integral_constant<int, 4>{});
  level(std::integral_constant<int, 5>{});

  // lane l holds ranks 64l .. 64l + 63: sixteen 16-byte stores of
  // consecutive ranks per lane (a partial run stores its first cnt ranks)
  const int o = RPL * lane;
  if (cnt == RUN) {
#pragma unroll
    for (int k = 0; k < RPL / 4; ++k) gstore4<DT, CONV_OUT>(dst, nullptr, base, o + 4 * k, &v[4 * k], flip);
  } else {
#pragma unroll
    for (int k = 0; k < RPL / 4; ++k) {
      if (o + 4 * k + 4 <= cnt) {
        gstore4<DT, CONV_OUT>(dst, nullptr, base, o + 4 * k, &v[4 * k], flip);
      } else {
#pragma unroll
        for (int t = 0; t < 3; ++t)
          if (o + 4 * k + t < cnt) gstore<DT, CONV_OUT>(dst, nullptr, base + o + 4 * k + t, v[4 * k + t], flip);
      }
    }
  }
}

// The plain sort's base case for 32-bit keys: runs of 4096 in place.
template <int DT, bool CONV_IN, bool CONV_OUT>
__global__ __launch_bounds__(BASE_WAVES * WAVE) void k_basecase_run(void* keys, int64_t n, uint32_t flip) {
  const int lane = threadIdx.x & (WAVE - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
  const int64_t base = ((int64_t)blockIdx.x * BASE_WAVES + wave) * RUN;
  if (base >= n) return;  // no barriers in this kernel
  const int64_t rest = n - base;
  sort_run<DT, CONV_IN, CONV_OUT>(keys, keys, base, rest < RUN ? (int)rest : RUN, flip, lane);
}

// Batched segments: run descriptors instead of one array. A descriptor holds
// `cnt` elements of one segment at `off` (1..TILE, or 1..4*TILE for the
// workgroup-run variant); it is read from buffer 0 in the caller's bits and
// written, sorted, to buffer (flags & 1), in the caller's bits when
// (flags & 2) (a segment of one run has no merge round).
struct TileDesc {
  int64_t off;
  int32_t cnt;
  int32_t flags;
};

// RUN4: one workgroup per descriptor of up to 4096 elements (four tiles of
// one segment), merged into one sorted run as in k_basecase; otherwise one
// wavefront per descriptor of up to 1024 elements.
template <int DT, bool RUN4 = false>
__global__ __launch_bounds__(BASE_WAVES* WAVE) void k_basecase_seg(void* k0, uint32_t* v0, void* k1, uint32_t* v1,
                                                                    const TileDesc* tiles, int64_t n_tiles,
                                                                    uint32_t flip) {
  using E = typename DType<DT>::E;
  __shared__ __align__(16) uint32_t xp[BASE_WAVES * E::W * XPOSE_WORDS];
  const int lane = threadIdx.x & (WAVE - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
  const int64_t tile = RUN4 ? (int64_t)blockIdx.x : (int64_t)blockIdx.x * BASE_WAVES + wave;
  if (tile >= n_tiles) return;  // RUN4: the whole workgroup leaves together
  const TileDesc td = tiles[tile];
  const int64_t base = RUN4 ? td.off + (int64_t)wave * TILE : td.off;
  const int rest = RUN4 ? td.cnt - wave * TILE : td.cnt;
  const int cnt = rest <= 0 ? 0 : (rest < TILE ? rest : TILE);

  E v[TPL];
  tile_load<DT, true>(k0, v0, base, cnt, flip, v, lane);

  sort_tile(v, lane);

  if constexpr (RUN4) {
    wave_exchange<true>(v, xp, wave, wave ^ 1, (wave & 1) == 0, lane);
    lane_cleaners<6>(v, lane);
    clean_regs(v);
    wave_exchange<true>(v, xp, wave, 3 - wave, wave < 2, lane);
    wave_exchange<false>(v, xp, wave, wave ^ 1, (wave & 1) == 0, lane);
    lane_cleaners<6>(v, lane);
    clean_regs(v);
    __syncthreads();
  }

  uint32_t* my = xp + wave * E::W * XPOSE_WORDS;
  void* dk = (td.flags & 1) ? k1 : k0;
  uint32_t* dv = (td.flags & 1) ? v1 : v0;
  if (td.flags & 2) tile_store<DT, true>(dk, dv, base, cnt, flip, v, my, lane);
  else tile_store<DT, false>(dk, dv, base, cnt, 0u, v, my, lane);
}

// Batched segments, 32-bit keys: one wavefront per run descriptor of up to
// 4096 elements (as k_basecase_seg with RUN4, without the four-wavefront
// exchanges).
template <int DT>
__global__ __launch_bounds__(BASE_WAVES* WAVE) void k_basecase_run_seg(void* k0, void* k1, const TileDesc* tiles,
                                                                        int64_t n_tiles, uint32_t flip) {
  const int lane = threadIdx.x & (WAVE - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
  const int64_t t = (int64_t)blockIdx.x * BASE_WAVES + wave;
  if (t >= n_tiles) return;  // no barriers in this kernel
  const TileDesc td = tiles[t];
  void* dk = (td.flags & 1) ? k1 : k0;
  if (td.flags & 2) sort_run<DT, true, true>(k0, dk, td.off, td.cnt, flip, lane);
  else sort_run<DT, true, false>(k0, dk, td.off, td.cnt, flip, lane);  // flip maps the load only
}

}  // namespace gmms
