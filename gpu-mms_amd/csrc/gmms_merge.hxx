// One merge round: partition (findPartitions) + K-way minBlockHeap merge
// (multimergeLevel / multimergePipeline).
//
// Reference: src/multimergesort.hxx:87-135 (round loop), :155-209
// (multimergeLevel), src/warp-partition.hxx:179-366,487-532 (findPartitions,
// single_pivot_partition), src/io-merge-gen.hxx:272-355 (buildHeap,
// multimergePipeline).
//
// A "task" is K consecutive sorted lists of length L (the last task may be
// ragged). A task of T elements is shared by m mergers; merger j writes output
// ranks [j*T/m, (j+1)*T/m) of the task (the reference's target
// warp-partition.hxx:234).
//
// Partition. Merger j needs a starting point. Instead of an exact co-rank at
// its target rank r, k_partition finds a *consistent cut*: per-list offsets
// c_k = number of elements of list k ordered before a pivot element y under
// the stable order (key, list, position). Every element before the cut
// precedes every element after it, so the merge of the lists from the cut on
// is exactly the task's sorted sequence from rank R = sum(c_k). The search
// keeps two consistent cuts lo (R_lo <= r) and hi (R_hi > r) and narrows them
// with one pivot per step: the weighted median (weights: the lists' window
// widths) of per-list candidates placed by interpolation for rank r, which
// halves r - R_lo or R_hi - r; after two such steps in a row that only lower
// R_hi, the candidates are the window midpoints instead (a median step,
// which removes >= 1/4 of the window) -- until r - R_lo <= slack. The heap
// merger then discards the first r - R_lo outputs (the tile mergers start at
// the cut itself). Each pivot needs one windowed search per list, so the
// cost is a few searches per list per merger; no list's pivot is ever
// written by an idle merger (the reference's race, warp-partition.hxx:522-531).
//
// Merge. A minBlockHeap of 2K-1 nodes x 64 elements (one element per lane)
// lives in LDS, word planes of 32 bits; every node's elements are <= all
// elements of its children, so the root holds the 64 smallest remaining
// elements. One wavefront per heap level (PIPELINE): each iteration the root
// is emitted, every level-wave merges the two children of its node on the
// root-to-leaf path (bitonic merge of 2x64 across lanes), keeps the lower 64
// and sends the upper 64 to the sibling of the next path node, and the
// emptied leaf is refilled from a per-list ring of prefetched blocks that
// gfx950's global_load_lds DMA writes straight into LDS. Internal nodes start
// as minimum-element blocks, so building the heap is just K-1 extra
// iterations whose emissions are discarded; termination is by count, never by
// a sentinel value (the reference stops at the first MAXVAL,
// io-merge-gen.hxx:324, and truncates real INT_MAX keys).
#pragma once
#include "gmms_elem.hxx"

namespace gmms {

constexpr int MAX_K = 64;
constexpr int MAX_PL = 6;

struct RoundArgs {
  const void* in_keys;
  const uint32_t* in_vals;
  void* out_keys;
  uint32_t* out_vals;
  int64_t n;      // elements in the array
  int64_t L;      // list length this round
  int64_t tasks;  // ceil(n / (K*L))
  int K;          // lists per task (<= 64)
  int PL;         // log2(K)
  int m;          // mergers per task
  int D;          // ring depth (blocks prefetched per list)
  int64_t slack;  // partition tolerance (elements discarded at most)
  int64_t* cuts;  // [(task*m + j) * cut_stride + k]
  int64_t* skips; // [task*m + j]
  int cut_stride; // entries per cut row: K (plain sort) or MAX_K (batched)
  uint32_t flip;  // XOR of the internal bits on the final store (descending)
  // batched segments: merger descriptors instead of the uniform task grid
  const struct MergerInfo* items;  // nullptr for a plain sort
  int64_t n_items;
  unsigned int* queue;             // work-queue counter of a persistent launch
  // every SAMP-th input element (internal bits), or nullptr: the partition
  // searches these first. Lists must start at multiples of SAMP.
  const void* samp_keys;
  const uint32_t* samp_vals;
  // where a merge writes the samples of its output for the next round's
  // partition (nullptr: none); the wavefront merger writes them
  void* samp_out_keys;
  uint32_t* samp_out_vals;
};

constexpr int SAMP_SHIFT = 8;
constexpr int64_t SAMP = (int64_t)1 << SAMP_SHIFT;  // sample stride (elements)

// Sample gather: samp[i] = in[i * SAMP] for i < ceil(n / SAMP).
template <int DT>
__global__ __launch_bounds__(256) void k_sample(const void* keys, const uint32_t* vals, int64_t n, void* sk,
                                                uint32_t* sv) {
  using Wd = typename DType<DT>::Word;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ((n + SAMP - 1) >> SAMP_SHIFT)) return;
  static_cast<Wd*>(sk)[i] = static_cast<const Wd*>(keys)[i << SAMP_SHIFT];
  if constexpr (DType<DT>::E::HAS_VAL) sv[i] = vals[i << SAMP_SHIFT];
}

// Geometry of one merger: task base (element index of its list 0), task size
// T, list length L, mergers per task m, its index j, and the cut row of its
// task's merger 0 (merger j reads cut rows cut_row+j and cut_row+j+1).
struct MergerInfo {
  int64_t tbase;
  int64_t T;
  int64_t L;
  int64_t cut_row;
  int32_t m;
  int32_t j;
  int32_t K;  // lists of this merger's task
  int32_t pad_;
};

__device__ __forceinline__ MergerInfo plain_merger(const RoundArgs& a, int64_t merger) {
  MergerInfo mi;
  const int64_t task = merger / a.m;
  mi.j = (int32_t)(merger % a.m);
  mi.m = a.m;
  mi.L = a.L;
  mi.tbase = task * (int64_t)a.K * a.L;
  const int64_t rest = a.n - mi.tbase;
  const int64_t full = (int64_t)a.K * a.L;
  mi.T = rest < full ? rest : full;
  mi.cut_row = task * a.m;
  mi.K = a.K;
  mi.pad_ = 0;
  return mi;
}
__device__ __forceinline__ MergerInfo load_merger(const MergerInfo* p) {
  MergerInfo mi;
  mi.tbase = p->tbase;
  mi.T = p->T;
  mi.L = p->L;
  mi.cut_row = p->cut_row;
  mi.m = p->m;
  mi.j = p->j;
  mi.K = p->K;
  mi.pad_ = 0;
  return mi;
}
__device__ __forceinline__ int64_t mlist_len(const MergerInfo& mi, int k) {
  int64_t l = mi.T - (int64_t)k * mi.L;
  if (l < 0) l = 0;
  return l < mi.L ? l : mi.L;
}

__device__ __forceinline__ int64_t task_size(const RoundArgs& a, int64_t task) {
  const int64_t tb = task * (int64_t)a.K * a.L;
  const int64_t rest = a.n - tb;
  const int64_t full = (int64_t)a.K * a.L;
  return rest < full ? rest : full;
}
__device__ __forceinline__ int64_t list_len(const RoundArgs& a, int64_t T, int k) {
  int64_t l = T - (int64_t)k * a.L;
  if (l < 0) l = 0;
  return l < a.L ? l : a.L;
}

template <int DT>
__device__ __forceinline__ typename DType<DT>::E ld_int(const RoundArgs& a, int64_t i) {
  return gload<DT, false>(a.in_keys, a.in_vals, i);
}
// input element i * SAMP, from the sample array
template <int DT>
__device__ __forceinline__ typename DType<DT>::E ld_samp(const RoundArgs& a, int64_t i) {
  return gload<DT, false>(a.samp_keys, a.samp_vals, i);
}

__device__ __forceinline__ int wave_sum32(int x) {
#pragma unroll
  for (int d = 1; d < WAVE; d <<= 1) x += __shfl_xor(x, d, WAVE);
  return x;
}

__device__ __forceinline__ int64_t wave_sum64(int64_t x) {
#pragma unroll
  for (int d = 1; d < WAVE; d <<= 1) x += __shfl_xor(x, d, WAVE);
  return x;
}

// -------------------------------------------------------------- partition
constexpr int PART_WAVES = 4;

// Consistent cut of merger mi.j >= 1 of its task: writes cut row
// mi.cut_row + mi.j (K offsets) and its skip.
template <int DT>
__device__ __forceinline__ void partition_one(const RoundArgs& a, const MergerInfo& mi, int lane) {
  using E = typename DType<DT>::E;
  const int j = mi.j;
  const int64_t T = mi.T;
  const int64_t r = (int64_t)j * T / mi.m;
  const int K = mi.K;
  const bool inl = lane < K;
  // the plain sort passes one slack per round; batched descriptors size it per task
  const int64_t slack = a.slack > 0 ? a.slack : max((int64_t)WAVE, T / mi.m / 32);
  const int64_t nk = inl ? mlist_len(mi, lane) : 0;
  const int64_t lbase = mi.tbase + (int64_t)lane * mi.L;

  int64_t lo = 0, hi = nk, Rlo = 0, Rhi = T;
  int hi_moves = 0;  // consecutive interpolation steps that only lowered hi
  for (int it = 0; it < 256 && r - Rlo > slack; ++it) {
    const int64_t Wt = Rhi - Rlo;  // == sum(hi - lo)
    const int64_t w = hi - lo;
    const bool act = w > 0;
    // candidates: each list's position interpolated for rank r, or its
    // midpoint (a median step, after two interpolation steps in a row that
    // only lowered hi); the pivot is their weighted median either way
    const bool interp = hi_moves < 2;
    const int64_t tw = Wt >> 1;  // in [0, Wt)
    int64_t pos = lo;
    bool from_samp = false;
    if (act) {
      int64_t off = interp ? (int64_t)((double)(r - Rlo) / (double)Wt * (double)w) : (w >> 1);
      if (off < 0) off = 0;
      pos = lo + (off < w ? off : w - 1);
      if (a.samp_keys) {
        // the nearest sample position at or above pos's block start inside
        // [lo, hi): read from the (cache-resident) sample array
        int64_t ps = pos & ~(SAMP - 1);
        if (ps < lo) ps += SAMP;
        if (ps < hi) {
          pos = ps;
          from_samp = true;
        }
      }
    }
    const E x = !act       ? Limits<E>::hi()
                : from_samp ? ld_samp<DT>(a, (lbase + pos) >> SAMP_SHIFT)
                            : ld_int<DT>(a, lbase + pos);
    // rank of this candidate among active candidates, order (element,
    // lane); the candidates are read with readlane (no LDS round trips)
    const unsigned long long amask = __ballot(act);
    int rank = 0;
    for (int i = 0; i < K; ++i) {
      if (!((amask >> i) & 1ull)) continue;  // wave-uniform
      uint32_t wi[E::W];
#pragma unroll
      for (int q = 0; q < E::W; ++q) wi[q] = (uint32_t)__builtin_amdgcn_readlane((int)word(x, q), i);
      const E xi = from_words<E>(wi);
      rank += (lt(xi, x) || (!lt(x, xi) && i < lane)) ? 1 : 0;
    }
    // inclusive weight prefix in sorted order
    int64_t pre = 0;
    for (int i = 0; i < K; ++i) {
      if (!((amask >> i) & 1ull)) continue;
      const int ri = __builtin_amdgcn_readlane(rank, i);
      const int64_t wi = (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)((uint64_t)w >> 32), i) << 32) |
                                   (uint32_t)__builtin_amdgcn_readlane((int)w, i));
      if (ri <= rank) pre += wi;
    }
    const bool is_piv = act && (pre - w) <= tw && tw < pre;
    const unsigned long long bal = __ballot(is_piv);
    const int kp = __ffsll((long long)bal) - 1;
    uint32_t yw[E::W];
#pragma unroll
    for (int q = 0; q < E::W; ++q) yw[q] = (uint32_t)__shfl((int)word(x, q), kp, WAVE);
    const E y = from_words<E>(yw);
    const int64_t ppos = __shfl(pos, kp, WAVE);
    // count of list elements ordered before y, inside [lo, hi]: every list
    // is searched by G = 64/K lanes at once: an A-ary search (A = np*G,
    // np = 1 probe per lane, 2 when G = 1) with probes g = t*G + gs
    // (g <= A-2) at l + (g+1)*n/A; powers of two, so shifts only. (More
    // probes per lane measured slower: the searches are bound by memory
    // transactions, not by round trips.)
    constexpr int PR = 2;
    const int G = WAVE / K;
    const int np = G >= 2 ? 1 : 2;
    const int ab = __builtin_ctz((unsigned)(G * np));  // log2(A)
    const int A = 1 << ab;
    const int gl = lane / G;  // this lane's list
    const int gs = lane % G;
    int64_t l = __shfl(lo, gl, WAVE), h = __shfl(hi, gl, WAVE);
    if (gl == kp) l = h = ppos;     // the pivot's own list: exactly its position
    const bool upper = gl < kp;     // equal keys of lower lists come first
    const int64_t gbase = mi.tbase + (int64_t)gl * mi.L;
    const unsigned long long gmask = G == WAVE ? ~0ull : ((1ull << G) - 1ull);
    // [sl, sh) -> the first index not before y (sh: none); ld(i) loads index i
    auto search = [&](int64_t& sl, int64_t& sh, auto ld) {
      for (;;) {
        const bool open = sl < sh;
        if (__ballot(open) == 0) break;
        const int64_t n = sh - sl;
        E v[PR];
#pragma unroll
        for (int t = 0; t < PR; ++t) {
          const int g = t * G + gs;
          if (t < np && open && g <= A - 2) v[t] = ld(sl + (((int64_t)(g + 1) * n) >> ab));
        }
        unsigned long long nb = 0;  // bit g: probe g is not before y
#pragma unroll
        for (int t = 0; t < PR; ++t) {
          if (t >= np) break;  // wave-uniform
          const int g = t * G + gs;
          const bool probe = open && g <= A - 2;
          const bool before = !probe || (upper ? !lt(y, v[t]) : lt(v[t], y));
          nb |= ((__ballot(!before) >> (gl * G)) & gmask) << (t * G);
        }
        if (open) {
          if (nb == 0) {
            sl = sl + (((int64_t)(A - 1) * n) >> ab) + 1;  // past the last probe
          } else {
            const int g0 = __ffsll((long long)nb) - 1;  // first probe not before y
            const int64_t hp = sl + (((int64_t)(g0 + 1) * n) >> ab);
            sl = g0 == 0 ? sl : sl + (((int64_t)g0 * n) >> ab) + 1;
            sh = hp;
          }
        }
      }
    };
    if (a.samp_keys) {
      // narrow [l, h) on the samples first: sample t sits at list position
      // t * SAMP; find the first sample in [l, h) not before y, then the
      // answer lies in the block that ends there
      const int64_t sl = (l + SAMP - 1) >> SAMP_SHIFT, sh = (h + SAMP - 1) >> SAMP_SHIFT;
      int64_t tl = sl, th = sh;
      const int64_t sbase = gbase >> SAMP_SHIFT;
      search(tl, th, [&](int64_t i) { return ld_samp<DT>(a, sbase + i); });
      if (tl > sl) l = max(l, ((tl - 1) << SAMP_SHIFT) + 1);
      if (tl < sh) h = min(h, tl << SAMP_SHIFT);
    }
    search(l, h, [&](int64_t i) { return ld_int<DT>(a, gbase + i); });
    int64_t c = __shfl(l, (lane % K) * G, WAVE);
    if (!inl) c = 0;
    const int64_t R = wave_sum64(c);
    if (R <= r) {
      lo = c;
      Rlo = R;
      if (R < r && lane == kp) lo = c + 1;  // the pivot itself is taken too
      if (R < r) Rlo = R + 1;
      hi_moves = 0;
    } else {
      hi = c;
      Rhi = R;
      hi_moves = interp ? hi_moves + 1 : 0;
    }
  }
  const int64_t b = mi.cut_row + j;
  if (inl) a.cuts[b * a.cut_stride + lane] = lo;
  if (lane == 0 && a.skips) a.skips[b] = r - Rlo;
}

// One wavefront per merger j >= 1: of the uniform task grid (plain sort), or
// of the descriptor list (batched segments; descriptors with j == 0 are not
// listed).
template <int DT>
__global__ __launch_bounds__(PART_WAVES * WAVE) void k_partition(RoundArgs a) {
  const int lane = threadIdx.x & (WAVE - 1);
  const int64_t wid = (int64_t)blockIdx.x * PART_WAVES + threadIdx.x / WAVE;
  if (a.items) {
    if (wid >= a.n_items) return;
    partition_one<DT>(a, load_merger(a.items + wid), lane);
    return;
  }
  const int64_t per_task = a.m - 1;
  if (wid >= a.tasks * per_task) return;
  const int64_t task = wid / per_task;
  partition_one<DT>(a, plain_merger(a, task * a.m + 1 + wid % per_task), lane);
}

// ------------------------------------------------------------------ merge
// s_waitcnt vmcnt(n) with n a run-time value: clamp and dispatch.
template <int N>
__device__ __forceinline__ void wait_vm_ge(int n) {
  if constexpr (N == 0) {
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0), expcnt/lgkmcnt untouched
  } else {
    if (n >= N) {
      // gfx9 encoding: vmcnt[3:0] bits 3:0, vmcnt[5:4] bits 15:14
      __builtin_amdgcn_s_waitcnt((N & 0xF) | ((N >> 4) << 14) | 0x0F70);
    } else {
      wait_vm_ge<N - 1>(n);
    }
  }
}
__device__ __forceinline__ void wait_vm(int n) { wait_vm_ge<31>(n); }

template <int D, class E>
__device__ __forceinline__ void clean64_step(E& v, int lane) {
  const bool low = (lane & D) == 0;
  v = lane_ce(v, xlane<D>(v), low, low ? 0u : 0xFFFFFFFFu);
}

template <class E>
__device__ __forceinline__ void bitonic_clean64(E& v, int lane) {
  clean64_step<32>(v, lane);
  clean64_step<16>(v, lane);
  clean64_step<8>(v, lane);
  clean64_step<4>(v, lane);
  clean64_step<2>(v, lane);
  clean64_step<1>(v, lane);
}

template <class E>
__device__ __forceinline__ E lds_get(const uint32_t* plane0, int stride, int idx) {
  uint32_t w[E::W];
#pragma unroll
  for (int q = 0; q < E::W; ++q) w[q] = plane0[q * stride + idx];
  return from_words<E>(w);
}
template <class E>
__device__ __forceinline__ void lds_put(uint32_t* plane0, int stride, int idx, const E& e) {
#pragma unroll
  for (int q = 0; q < E::W; ++q) plane0[q * stride + idx] = word(e, q);
}

// LDS words needed by one merger
__host__ __device__ inline size_t merge_lds_words(int K, int W, int D) {
  const size_t nodes = (size_t)(2 * K - 1);
  return nodes * W * WAVE + nodes * W + (size_t)K * D * W * WAVE + (size_t)K * D;
}

template <int DT, bool CONV_OUT>
__global__ __launch_bounds__(MAX_PL * WAVE) void k_merge(RoundArgs a) {
  using E = typename DType<DT>::E;
  constexpr int W = E::W;
  extern __shared__ uint32_t lds[];
  const int K = a.K, PL = a.PL, D = a.D;
  const int NN = 2 * K - 1;
  uint32_t* nodes = lds;                            // [node][plane][64]
  uint32_t* nmax = nodes + (size_t)NN * W * WAVE;   // [plane][node]
  uint32_t* ring = nmax + (size_t)NN * W;           // [k][d][plane][64]
  int* issued = (int*)(ring + (size_t)K * D * W * WAVE);  // [k*D + d]

  const int lane = threadIdx.x & (WAVE - 1);
  const int wave = threadIdx.x / WAVE;
  const int64_t task = blockIdx.x / a.m;
  const int j = blockIdx.x % a.m;
  const int64_t T = task_size(a, task);
  const int64_t tbase = task * (int64_t)K * a.L;
  const int64_t r0 = (int64_t)j * T / a.m;
  const int64_t r1 = (int64_t)(j + 1) * T / a.m;
  const int64_t count = r1 - r0;
  const int64_t skip = j == 0 ? 0 : a.skips[blockIdx.x];

  // internal nodes start as minimum blocks
  for (int i = threadIdx.x; i < (K - 1) * W * WAVE; i += blockDim.x) nodes[i] = 0u;
  for (int i = threadIdx.x; i < NN * W; i += blockDim.x) nmax[i] = 0u;
  __syncthreads();

  const bool bottom = wave == PL - 1;
  int64_t lstart = 0, lrem = 0;  // lane k: list k from its cut
  int cons = 1;                  // lane k: next block index to consume
  int G = 0;                     // DMA instructions issued by the bottom wave

  auto dma_block = [&](int k, int64_t st, int64_t rem, int blk, uint32_t* dst) {
    const int64_t idx = (int64_t)blk * WAVE + lane;
    if (idx < rem) {
      const int64_t gi = st + idx;
      if constexpr (W == 1) {
        __builtin_amdgcn_global_load_lds((const void*)((const uint32_t*)a.in_keys + gi),
                                         (__attribute__((address_space(3))) void*)dst, 4, 0, 0);
      } else {
        const uint32_t* kw = (const uint32_t*)a.in_keys + 2 * gi;
        __builtin_amdgcn_global_load_lds((const void*)kw, (__attribute__((address_space(3))) void*)dst, 4, 0, 0);
        __builtin_amdgcn_global_load_lds((const void*)(kw + 1),
                                         (__attribute__((address_space(3))) void*)(dst + WAVE), 4, 0, 0);
        if constexpr (W == 3) {
          __builtin_amdgcn_global_load_lds((const void*)(a.in_vals + gi),
                                           (__attribute__((address_space(3))) void*)(dst + 2 * WAVE), 4, 0, 0);
        }
      }
    }
    (void)k;
  };

  if (bottom) {
    if (lane < K) {
      const int64_t nk = list_len(a, T, lane);
      const int64_t ck = j == 0 ? 0 : a.cuts[(int64_t)blockIdx.x * K + lane];
      lstart = tbase + (int64_t)lane * a.L + ck;
      lrem = nk - ck;
    }
    for (int k = 0; k < K; ++k) {
      const int64_t st = __shfl(lstart, k, WAVE);
      const int64_t rem = __shfl(lrem, k, WAVE);
      uint32_t* leaf = nodes + (size_t)(K - 1 + k) * W * WAVE;
      if ((int64_t)lane >= rem) lds_put<E>(leaf, WAVE, lane, Limits<E>::hi());
      dma_block(k, st, rem, 0, leaf);
      for (int d = 0; d < D; ++d) {
        dma_block(k, st, rem, d + 1, ring + ((size_t)(k * D + d) * W) * WAVE);
        G += W;
        if (lane == 0) issued[k * D + d] = G;
      }
    }
    wait_vm(0);
    __builtin_amdgcn_s_waitcnt(0);
    if (lane < K) {
      const E mx = lrem >= WAVE ? lds_get<E>(nodes + (size_t)(K - 1 + lane) * W * WAVE + (WAVE - 1), WAVE, 0)
                                : Limits<E>::hi();
      lds_put<E>(nmax, NN, K - 1 + lane, mx);
    }
  }
  __syncthreads();

  E root = Limits<E>::lo();
  const int64_t pre = (int64_t)WAVE * (K - 1) + skip;  // discarded emissions
  const int64_t iters = (pre + count + WAVE - 1) / WAVE;
  const int64_t obase = tbase + r0;
  for (int64_t it = 0; it < iters; ++it) {
    if (wave == 0) {
      const int64_t o = it * WAVE + lane - pre;
      if (o >= 0 && o < count) gstore<DT, CONV_OUT>(a.out_keys, a.out_vals, obase + o, root, a.flip);
    }
    if (it == iters - 1) break;
    // root-to-leaf path: the child with the smaller maximum is emptied
    int p = 0, pn = 0;
    for (int l = 0; l <= wave; ++l) {
      const int c1 = 2 * p + 1;
      const E m1 = lds_get<E>(nmax, NN, c1);
      const E m2 = lds_get<E>(nmax, NN, c1 + 1);
      const int q = le(m1, m2) ? c1 : c1 + 1;
      if (l < wave) p = q; else pn = q;
    }
    const int c1 = 2 * p + 1;
    const E va = lds_get<E>(nodes + (size_t)c1 * W * WAVE, WAVE, lane);
    const E vb = lds_get<E>(nodes + (size_t)(c1 + 1) * W * WAVE, WAVE, WAVE - 1 - lane);
    __syncthreads();
    E lo = emin(va, vb), hi = emax(va, vb);
    bitonic_clean64(lo, lane);
    bitonic_clean64(hi, lane);
    const int sib = (pn == c1) ? c1 + 1 : c1;
    lds_put<E>(nodes + (size_t)sib * W * WAVE, WAVE, lane, hi);
    if (lane == WAVE - 1) lds_put<E>(nmax, NN, sib, hi);
    if (wave == 0) {
      root = lo;
    } else {
      lds_put<E>(nodes + (size_t)p * W * WAVE, WAVE, lane, lo);
      if (lane == WAVE - 1) lds_put<E>(nmax, NN, p, lo);
    }
    if (bottom) {
      // refill leaf pn from the ring
      const int k = pn - (K - 1);
      const int blk = __shfl(cons, k, WAVE);
      const int64_t st = __shfl(lstart, k, WAVE);
      const int64_t rem = __shfl(lrem, k, WAVE);
      const int d = (blk - 1) % D;
      const int isd = issued[k * D + d];
      if (PL == 1) wait_vm(0); else wait_vm(G - isd);
      uint32_t* slot = ring + ((size_t)(k * D + d) * W) * WAVE;
      const int64_t idx = (int64_t)blk * WAVE + lane;
      E v = idx < rem ? lds_get<E>(slot, WAVE, lane) : Limits<E>::hi();
      uint32_t* leaf = nodes + (size_t)pn * W * WAVE;
      lds_put<E>(leaf, WAVE, lane, v);
      if (lane == WAVE - 1) lds_put<E>(nmax, NN, pn, v);
      if (lane == k) cons = blk + 1;
      if ((int64_t)(blk + D) * WAVE < rem) {
        __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0): slot read done before DMA reuses it
        dma_block(k, st, rem, blk + D, slot);
        G += W;
        if (lane == 0) issued[k * D + d] = G;
      }
    }
    __syncthreads();
  }
}

}  // namespace gmms
