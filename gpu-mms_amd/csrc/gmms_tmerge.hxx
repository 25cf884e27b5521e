// Tile K-way merge: the second merge kernel of a round, for the same
// partition (consistent cuts from k_partition) as the minBlockHeap kernel.
//
// Replaces the per-merger work of multimergeLevel/multimergePipeline
// (src/multimergesort.hxx:155-209, src/io-merge-gen.hxx:305-355) with a
// design whose cost per element is O(log K) cheap LDS steps instead of
// O(log K * log 64) cross-lane exchange stages:
//
// A workgroup owns the elements between two consistent cuts (lo = cut of
// merger j, hi = cut of merger j+1) and writes them at rank sum(lo). It keeps
// an LDS window of the next F elements of every list and repeats:
//   1. refill every window to min(F, remaining) elements;
//   2. frontier y = the smallest "last window element" over the lists whose
//      window is full (tag order (key, list)); every remaining element <= y
//      is already in a window, so taking exactly those keeps the cut
//      consistent and takes >= F elements;
//   3. if only one list has elements <= y (sorted or clustered input), gallop
//      instead: the whole run of that list below the next list's first
//      element is the next output, copied global -> global;
//   4. otherwise gather the K taken prefixes into one LDS buffer and merge
//      them with log2 K passes of stable pairwise merge path (each thread: one
//      diagonal binary search + a short serial merge), then store coalesced.
// Stability inside the merge (left run first on ties) plus the (key, list)
// frontier make every step's cut consistent, so the output equals the
// stable K-way merge bit for bit.
#pragma once
#include "gmms_elem.hxx"
#include "gmms_merge.hxx"

namespace gmms {

constexpr int TM_THREADS = 256;
constexpr int TM_WAVES = TM_THREADS / WAVE;

// LDS words of one tile merger: window + two merge buffers (K*F elements
// each, W word planes), then int64 cur/end per list + int64 scalars, then
// int head/cnt/want/taken per list, off[K+1] and int scalars.
__host__ __device__ inline size_t tmerge_lds_words(int K, int W, int F) {
  const size_t cap = (size_t)K * F;
  return 3 * cap * W + 2 * (2 * (size_t)K + 4) + 4 * (size_t)K + (K + 1) + 8;
}

template <class E>
__device__ __forceinline__ E pl_get(const uint32_t* base, size_t cap, int idx) {
  uint32_t w[E::W];
#pragma unroll
  for (int q = 0; q < E::W; ++q) w[q] = base[q * cap + idx];
  return from_words<E>(w);
}
template <class E>
__device__ __forceinline__ void pl_put(uint32_t* base, size_t cap, int idx, const E& e) {
#pragma unroll
  for (int q = 0; q < E::W; ++q) base[q * cap + idx] = word(e, q);
}

// tag order (element, list): does (v, kv) precede (y, ky)?
template <class E>
__device__ __forceinline__ bool tag_lt(const E& v, int kv, const E& y, int ky) {
  return lt(v, y) || (!lt(y, v) && kv < ky);
}

template <int DT, bool CONV_OUT, int F>
__global__ __launch_bounds__(TM_THREADS) void k_tmerge(RoundArgs a) {
  using E = typename DType<DT>::E;
  constexpr int W = E::W;
  extern __shared__ uint32_t lds[];
  const int K = a.K;
  const size_t cap = (size_t)K * F;
  uint32_t* win = lds;
  uint32_t* bufA = win + cap * W;
  uint32_t* bufB = bufA + cap * W;
  static_assert((F & (F - 1)) == 0 && F >= WAVE, "F: power of two >= 64");
  int64_t* cur = (int64_t*)(bufB + cap * W);  // 8-byte aligned: 3*cap*W words is even
  int64_t* endp = cur + K;
  int64_t* sh64 = endp + K;  // [0] output base, [1] gallop length
  int* cnt = (int*)(sh64 + 4);
  int* head = cnt + K;
  int* want = head + K;
  int* taken = want + K;
  int* off = taken + K;      // K + 1 entries
  int* sh = off + (K + 1);   // [0] S, [1] mode (0 merge, 1 gallop, 2 done), [2] gallop list
  int64_t* obase_p = sh64;

  const int tid = threadIdx.x;
  const int lane = tid & (WAVE - 1);
  const int wave = tid / WAVE;
  const int64_t task = blockIdx.x / a.m;
  const int j = blockIdx.x % a.m;
  const int64_t T = task_size(a, task);
  const int64_t tbase = task * (int64_t)K * a.L;

  if (wave == 0) {
    int64_t lo = 0, hi = 0, nk = 0;
    if (lane < K) {
      nk = list_len(a, T, lane);
      lo = j == 0 ? 0 : a.cuts[(int64_t)blockIdx.x * K + lane];
      hi = j == a.m - 1 ? nk : a.cuts[((int64_t)blockIdx.x + 1) * K + lane];
      const int64_t lb = tbase + (int64_t)lane * a.L;
      cur[lane] = lb + lo;
      endp[lane] = lb + hi;
      cnt[lane] = 0;
      head[lane] = 0;
    }
    const int64_t R = wave_sum64(lo);
    if (lane == 0) *obase_p = tbase + R;
  }
  __syncthreads();

  for (;;) {
    // ---- 1. refill windows (wave 0 first publishes want[])
    if (wave == 0 && lane < K) {
      const int64_t rem = endp[lane] - cur[lane];
      want[lane] = (int)(rem < F ? rem : F);
    }
    __syncthreads();
    {
      // every load of this wave in flight before the first LDS store
      // unconditional loads (element 0 when the slot needs nothing): no
      // branch join between them, so they are all in flight together
      constexpr int PER = MAX_K / TM_WAVES * (F / WAVE);
      E v[PER];
      uint32_t vmask = 0;
#pragma unroll
      for (int u = 0; u < PER; ++u) {
        const int k = wave + (u / (F / WAVE)) * TM_WAVES;
        const int i = (u % (F / WAVE)) * WAVE + lane;
        const int kk = k < K ? k : 0;
        const bool ok = k < K && i >= cnt[kk] && i < want[kk];
        v[u] = gload<DT, false>(a.in_keys, a.in_vals, ok ? cur[kk] + i : 0);
        vmask |= ok ? (1u << u) : 0u;
      }
#pragma unroll
      for (int u = 0; u < PER; ++u) {
        const int k = wave + (u / (F / WAVE)) * TM_WAVES;
        const int i = (u % (F / WAVE)) * WAVE + lane;
        if ((vmask >> u) & 1u) pl_put<E>(win, cap, k * F + ((head[k] + i) & (F - 1)), v[u]);
      }
    }
    __syncthreads();

    // ---- 2./3. frontier, single-list test, counts (wave 0, lane = list)
    if (wave == 0) {
      const bool inl = lane < K;
      const int wk = inl ? want[lane] : 0;
      const int hk = inl ? head[lane] : 0;
      const bool full = inl && wk == F;
      E last = full ? pl_get<E>(win, cap, lane * F + ((hk + F - 1) & (F - 1))) : Limits<E>::hi();
      // argmin over full lanes of (last, lane)
      E y = last;
      int ky = full ? lane : WAVE;
#pragma unroll
      for (int d = 1; d < WAVE; d <<= 1) {
        uint32_t w[W];
#pragma unroll
        for (int q = 0; q < W; ++q) w[q] = (uint32_t)__shfl_xor((int)word(y, q), d, WAVE);
        const E oy = from_words<E>(w);
        const int oky = __shfl_xor(ky, d, WAVE);
        const bool take = oky < WAVE && (ky == WAVE || tag_lt(oy, oky, y, ky));
        if (take) {
          y = oy;
          ky = oky;
        }
      }
      const bool anyfull = ky < WAVE;
      int mode = 0, sk = 0;
      if (!anyfull) {
        sk = wk;  // everything left fits in the windows: take it all
        const int tot = wave_sum32(sk);
        mode = tot == 0 ? 2 : 0;
      } else {
        // z = min over the other non-empty lists of their first element
        const bool nonempty = inl && wk > 0 && lane != ky;
        E z = nonempty ? pl_get<E>(win, cap, lane * F + hk) : Limits<E>::hi();
        int kz = nonempty ? lane : WAVE;
#pragma unroll
        for (int d = 1; d < WAVE; d <<= 1) {
          uint32_t w[W];
#pragma unroll
          for (int q = 0; q < W; ++q) w[q] = (uint32_t)__shfl_xor((int)word(z, q), d, WAVE);
          const E oz = from_words<E>(w);
          const int okz = __shfl_xor(kz, d, WAVE);
          const bool take = okz < WAVE && (kz == WAVE || tag_lt(oz, okz, z, kz));
          if (take) {
            z = oz;
            kz = okz;
          }
        }
        const bool single = kz == WAVE || tag_lt(y, ky, z, kz);
        if (single) {
          mode = 1;
          // gallop list ky from cur + F: first position whose element is not before z
          int64_t lo = F, hi = endp[ky] - cur[ky];  // answer in [lo, hi]
          const int64_t base = cur[ky];
          if (kz != WAVE) {
            while (hi > lo) {
              const int64_t span = hi - lo;
              const int64_t step = (span + WAVE - 1) / WAVE;
              const int64_t p = lo + (int64_t)lane * step;  // probe p: element at base + p
              bool before = true;
              if (p < hi) {
                const E v = gload<DT, false>(a.in_keys, a.in_vals, base + p);
                before = tag_lt(v, ky, z, kz);
              }
              const unsigned long long nb = __ballot(!before && p < hi);
              if (nb == 0) {
                // every probe is before z: the answer is past the last probe
                const int64_t nprobe = (hi - 1 - lo) / step;  // index of the last probe < hi
                const int64_t lastp = lo + (nprobe < WAVE - 1 ? nprobe : WAVE - 1) * step;
                lo = lastp + 1;
              } else {
                const int l0 = __ffsll((long long)nb) - 1;  // first lane not before z
                const int64_t phi = lo + (int64_t)l0 * step;
                const int64_t plo = l0 == 0 ? lo : lo + (int64_t)(l0 - 1) * step + 1;
                lo = plo;
                hi = phi;
              }
            }
          } else {
            lo = hi;
          }
          if (lane == 0) {
            sh[2] = ky;
            sh64[1] = lo;  // run length
          }
        } else {
          // count window elements <= (y, ky) in tag order
          if (inl && wk > 0) {
            if (lane == ky) {
              sk = F;
            } else {
              int l = 0, h = wk;
              while (l < h) {
                const int mid = (l + h) >> 1;
                const E v = pl_get<E>(win, cap, lane * F + ((hk + mid) & (F - 1)));
                const bool take = lane < ky ? !lt(y, v) : lt(v, y);
                if (take) l = mid + 1; else h = mid;
              }
              sk = l;
            }
          }
        }
      }
      if (mode == 0) {
        // exclusive scan of sk -> off
        int incl = sk;
#pragma unroll
        for (int d = 1; d < WAVE; d <<= 1) {
          const int o = __shfl_up(incl, d, WAVE);
          if (lane >= d) incl += o;
        }
        if (lane < K) off[lane] = incl - sk;
        if (lane == WAVE - 1) {
          off[K] = incl;
          sh[0] = incl;
        }
        if (lane < K) taken[lane] = sk;
      }
      if (lane == 0) sh[1] = mode;
    }
    __syncthreads();
    const int mode = sh[1];
    if (mode == 2) break;

    if (mode == 1) {
      // ---- gallop: copy the run of list kstar straight to the output
      const int ks = sh[2];
      const int64_t len = sh64[1];
      const int64_t src = cur[ks];
      const int64_t dst = *obase_p;
      for (int64_t i = tid; i < len; i += TM_THREADS) {
        const E v = gload<DT, false>(a.in_keys, a.in_vals, src + i);
        gstore<DT, CONV_OUT>(a.out_keys, a.out_vals, dst + i, v, a.flip);
      }
      __syncthreads();
      if (tid == 0) {
        cur[ks] += len;
        cnt[ks] = 0;  // window restarts at the new cursor
        head[ks] = 0;
        *obase_p += len;
      }
      // other lists keep their windows
      __syncthreads();
      continue;
    }

    // ---- 4. gather taken prefixes into bufA (runs at off[k])
    const int S = sh[0];
    for (int k = wave; k < K; k += TM_WAVES) {
      const int sk = taken[k], ok = off[k], hk = head[k];
      for (int i = lane; i < sk; i += WAVE) pl_put<E>(bufA, cap, ok + i, pl_get<E>(win, cap, k * F + ((hk + i) & (F - 1))));
    }
    __syncthreads();

    // ---- merge passes: groups of 2^(p+1) runs
    uint32_t* src = bufA;
    uint32_t* dst = bufB;
    const int E_per = (S + TM_THREADS - 1) / TM_THREADS;
    for (int g = 2; g <= K; g <<= 1) {
      int o = tid * E_per;
      const int oend = min(S, o + E_per);
      while (o < oend) {
        // group containing o: largest q with off[q*g] <= o
        int l = 0, h = K / g - 1;
        while (l < h) {
          const int mid = (l + h + 1) >> 1;
          const bool c = off[mid * g] <= o;
          l = c ? mid : l;
          h = c ? h : mid - 1;
        }
        const int gs = off[l * g], gm = off[l * g + g / 2], ge = off[l * g + g];
        const int n1 = gm - gs, n2 = ge - gm;
        const int d = o - gs;
        int lo = d - n2 > 0 ? d - n2 : 0;
        int hi = d < n1 ? d : n1;
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          // A1[mid] is among the first d outputs iff A1[mid] <= A2[d-1-mid] (left run first on ties)
          const bool c = !lt(pl_get<E>(src, cap, gm + d - 1 - mid), pl_get<E>(src, cap, gs + mid));
          lo = c ? mid + 1 : lo;
          hi = c ? hi : mid;
        }
        int i = lo, jj = d - lo;
        const int cntm = min(oend, ge) - o;
        // Reads past a run's end land on the next run or on the LDS words
        // after the buffer (index <= cap): in bounds, and never selected.
        E va = pl_get<E>(src, cap, gs + i);
        E vb = pl_get<E>(src, cap, gm + jj);
        for (int e = 0; e < cntm; ++e) {
          const bool ta = (jj >= n2) | ((i < n1) & !lt(vb, va));
          pl_put<E>(dst, cap, o + e, ta ? va : vb);
          i += ta ? 1 : 0;
          jj += ta ? 0 : 1;
          const E nv = pl_get<E>(src, cap, ta ? gs + i : gm + jj);
          va = ta ? nv : va;
          vb = ta ? vb : nv;
        }
        o += cntm;
      }
      __syncthreads();
      uint32_t* t = src;
      src = dst;
      dst = t;
    }

    // ---- store the merged step
    const int64_t ob = *obase_p;
    for (int o = tid; o < S; o += TM_THREADS) gstore<DT, CONV_OUT>(a.out_keys, a.out_vals, ob + o, pl_get<E>(src, cap, o), a.flip);
    __syncthreads();
    if (wave == 0) {
      if (lane < K) {
        const int sk = taken[lane];
        cur[lane] += sk;
        head[lane] = (head[lane] + sk) & (F - 1);
        cnt[lane] = want[lane] - sk;
      }
      if (lane == 0) *obase_p = ob + S;
    }
    __syncthreads();
  }
}

}  // namespace gmms
