// Element types, key-order maps and 64-lane cross-lane exchanges for gfx950.
//
// Every kernel works on an "internal" element whose order is plain unsigned
// order: the comparator contract of the reference (src/cmp.hxx:37-54,
// cmp(a,b) == "a strictly precedes b", MAXVAL larger than every key) is met
// by mapping each user dtype to unsigned bits once, on the base-case load, and
// back once, on the last round's store:
//   int32/int64  : flip the sign bit
//   float32/64   : IEEE total order (negatives reversed, -0 before +0)
//   (u64 key, u32 value) pairs : lexicographic (key, value), no map
// so ties are broken by the full element bits and the sorted output is unique.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gmms {

constexpr int WAVE = 64;

// dtype codes; must match include/gpu_mms/gmms.h
enum : int {
  DT_I32 = 0, DT_U32 = 1, DT_F32 = 2, DT_I64 = 3, DT_U64 = 4, DT_F64 = 5, DT_PAIR = 6
};

// ------------------------------------------------------------------ words
// Cross-lane value of lane (lane ^ M) for one 32-bit word. Masks below 64
// use DPP (quad_perm, row mirrors, row_ror:8) or ds_swizzle inside 32 lanes;
// 32 and 63 cross the half-wave and go through ds_bpermute.
// every lane of these patterns reads a valid source lane, so no old value
// needs to be carried (bound_ctrl set, no copy into the destination first)
template <int CTRL>
__device__ __forceinline__ uint32_t dpp_mov(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, CTRL, 0xF, 0xF, true);
}
template <int PAT>
__device__ __forceinline__ uint32_t swizzle(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_ds_swizzle((int)x, PAT);
}
template <int M>
__device__ __forceinline__ uint32_t xor_lane(uint32_t x) {
  static_assert(M > 0 && M < 64, "lane mask");
  if constexpr (M == 1) return dpp_mov<0xB1>(x);        // quad_perm [1,0,3,2]
  else if constexpr (M == 2) return dpp_mov<0x4E>(x);   // quad_perm [2,3,0,1]
  else if constexpr (M == 3) return dpp_mov<0x1B>(x);   // quad_perm [3,2,1,0]
  else if constexpr (M == 7) return dpp_mov<0x141>(x);  // row_half_mirror
  else if constexpr (M == 8) return dpp_mov<0x128>(x);  // row_ror:8
  else if constexpr (M == 15) return dpp_mov<0x140>(x); // row_mirror
  else if constexpr (M < 32) return swizzle<0x1F | (M << 10)>(x);  // bit-mode xor
  else return (uint32_t)__shfl_xor((int)x, M, 64);
}

// a wave-uniform 64-bit value, held in scalar registers
__device__ __forceinline__ int64_t uniform64(int64_t x) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)x);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)((uint64_t)x >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// Chunk swizzle of the register-tile transpose (XC = 4 chunks of 4 words
// per lane row of 16 words): a lane's chunk k goes to chunk position
// k ^ xswz(lane). Lanes 0,2,4,6 (and 1,3,5,7) get 0,1,3,2, and lanes m, m+4,
// m+8, m+12 get four distinct values, so the 16-byte accesses of 8
// consecutive lanes (32 banks) or 16 consecutive lanes (64 banks) touch every
// bank once.
template <int XC>
__device__ __forceinline__ int xswz(int l) {
  static_assert(XC == 4, "rows of 16 words");
  return ((l >> 1) ^ (l >> 2)) & 3;
}

// ---------------------------------------------------------------- elements
struct E32 {
  static constexpr int W = 1;
  static constexpr bool HAS_VAL = false;
  uint32_t k;
};
struct E64 {
  static constexpr int W = 2;
  static constexpr bool HAS_VAL = false;
  uint64_t k;
};
struct P96 {
  static constexpr int W = 3;
  static constexpr bool HAS_VAL = true;
  uint64_t k;
  uint32_t v;
};

// The (u64 key, u32 value) order as a lane mask: bit l is lane l's a < b,
// the borrow of a - b over (key high, key low, value) from three 32-bit
// subtractions chained through the carry, instead of two 64-bit compares
// (v_cmp_*_u64, several times the cost of a 32-bit instruction on gfx950), a
// 32-bit compare and two mask operations. The compare-exchanges of pairs use
// it directly (2^27 pairs 7.99 -> 7.49 ms). (For 64-bit keys alone, and for
// the compares that steer control flow, the compiler's own compares measured
// faster in the kernels.) Every output of the asm block is early-clobber:
// the block writes its first temporaries before it reads its last inputs,
// so no output may share a register with an input.
// (u64 key, u32 value) pairs: the borrow over (key high, key low, value)
__device__ __forceinline__ uint64_t lt_mask96(uint64_t ak, uint32_t av, uint64_t bk, uint32_t bv) {
  uint32_t t0, t1, t2;
  uint64_t m;
  asm("v_sub_co_u32 %0, %3, %4, %5\n\t"
      "v_subb_co_u32 %1, %3, %6, %7, %3\n\t"
      "v_subb_co_u32 %2, %3, %8, %9, %3\n\t"
      "s_nop 1"
      : "=&v"(t0), "=&v"(t1), "=&v"(t2), "=&s"(m)
      : "v"(av), "v"(bv), "v"((uint32_t)ak), "v"((uint32_t)bk), "v"((uint32_t)(ak >> 32)),
        "v"((uint32_t)(bk >> 32)));
  (void)t0;
  (void)t1;
  (void)t2;
  return m;
}

// Element compares for control flow (searches, the frontier, gallops) stay
// plain C++: there the masks above measured slower (the asm blocks keep the
// compiler from scheduling around them); they pay in the compare-exchanges,
// which use them directly (cswap, lane_ce below).
__device__ __forceinline__ bool lt(const E32& a, const E32& b) { return a.k < b.k; }
__device__ __forceinline__ bool lt(const E64& a, const E64& b) { return a.k < b.k; }
__device__ __forceinline__ bool lt(const P96& a, const P96& b) {
  return a.k < b.k || (a.k == b.k && a.v < b.v);
}
__device__ __forceinline__ bool le(const E32& a, const E32& b) { return a.k <= b.k; }
__device__ __forceinline__ bool le(const E64& a, const E64& b) { return a.k <= b.k; }
__device__ __forceinline__ bool le(const P96& a, const P96& b) { return !lt(b, a); }

// c ? a : b field by field (a select of a whole padded struct can end up in
// scratch memory)
__device__ __forceinline__ E32 sel(bool c, const E32& a, const E32& b) { return {c ? a.k : b.k}; }
__device__ __forceinline__ E64 sel(bool c, const E64& a, const E64& b) { return {c ? a.k : b.k}; }
__device__ __forceinline__ P96 sel(bool c, const P96& a, const P96& b) { return {c ? a.k : b.k, c ? a.v : b.v}; }

template <class E> __device__ __forceinline__ E emin(const E& a, const E& b) { return sel(lt(b, a), b, a); }
template <class E> __device__ __forceinline__ E emax(const E& a, const E& b) { return sel(lt(b, a), a, b); }
template <> __device__ __forceinline__ E32 emin(const E32& a, const E32& b) { return {a.k < b.k ? a.k : b.k}; }
template <> __device__ __forceinline__ E32 emax(const E32& a, const E32& b) { return {a.k < b.k ? b.k : a.k}; }

// compare-exchange: a <- min, b <- max (one comparison for multi-word
// elements; min/max instructions for 32-bit ones)
template <class E> __device__ __forceinline__ void cswap(E& a, E& b) {
  const bool c = lt(b, a);
  const E lo = sel(c, b, a), hi = sel(c, a, b);
  a = lo;
  b = hi;
}
template <> __device__ __forceinline__ void cswap(E32& a, E32& b) {
  const E32 lo = emin(a, b), hi = emax(a, b);
  a = lo;
  b = hi;
}
template <> __device__ __forceinline__ void cswap(P96& a, P96& b);  // with the lane-mask helpers below

// Cross-lane compare-exchange, one side: the lower lane of a pair keeps
// min(x, p), the upper keeps max(x, p). hsel is 0 on lower lanes and all
// ones on upper lanes: for 32-bit elements med3(x, p, hsel) is exactly that
// (one v_med3_u32 instead of min, max and a select).
__device__ __forceinline__ uint32_t med3_u32(uint32_t a, uint32_t b, uint32_t c) {
  uint32_t r;
  asm("v_med3_u32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
  return r;
}
template <class E>
__device__ __forceinline__ E lane_ce(const E& x, const E& p, bool low, uint32_t hsel) {
  (void)hsel;
  // one comparison: the lower lane takes the partner if it is smaller, the
  // upper lane if it is not (equal elements are identical either way)
  const bool c = lt(p, x);
  return sel(low == c, p, x);
}
template <>
__device__ __forceinline__ E32 lane_ce(const E32& x, const E32& p, bool low, uint32_t hsel) {
  (void)low;
  return {med3_u32(x.k, p.k, hsel)};
}

// per lane: bit l of m ? t : f
__device__ __forceinline__ uint32_t mask_sel(uint64_t m, uint32_t t, uint32_t f) {
  uint32_t r;
  asm("v_cndmask_b32 %0, %1, %2, %3" : "=v"(r) : "v"(f), "v"(t), "s"(m));
  return r;
}
__device__ __forceinline__ P96 mask_sel(uint64_t m, const P96& t, const P96& f) {
  const uint32_t lo = mask_sel(m, (uint32_t)t.k, (uint32_t)f.k);
  const uint32_t hi = mask_sel(m, (uint32_t)(t.k >> 32), (uint32_t)(f.k >> 32));
  return {((uint64_t)hi << 32) | lo, mask_sel(m, t.v, f.v)};
}
__device__ __forceinline__ uint64_t lt_mask(const P96& a, const P96& b) { return lt_mask96(a.k, a.v, b.k, b.v); }
// lower lanes take p where p < x, upper lanes where p >= x
template <>
__device__ __forceinline__ P96 lane_ce(const P96& x, const P96& p, bool low, uint32_t hsel) {
  (void)hsel;
  return mask_sel(~(lt_mask(p, x) ^ (uint64_t)__ballot(low)), p, x);
}
template <> __device__ __forceinline__ void cswap(P96& a, P96& b) {
  const uint64_t c = lt_mask(b, a);
  const P96 lo = mask_sel(c, b, a), hi = mask_sel(c, a, b);
  a = lo;
  b = hi;
}


template <class E> struct Limits;
template <> struct Limits<E32> {
  __device__ static E32 lo() { return {0u}; }
  __device__ static E32 hi() { return {0xFFFFFFFFu}; }
};
template <> struct Limits<E64> {
  __device__ static E64 lo() { return {0ull}; }
  __device__ static E64 hi() { return {~0ull}; }
};
template <> struct Limits<P96> {
  __device__ static P96 lo() { return {0ull, 0u}; }
  __device__ static P96 hi() { return {~0ull, 0xFFFFFFFFu}; }
};

// 32-bit word w of an element (word planes in LDS, cross-lane moves)
__device__ __forceinline__ uint32_t word(const E32& e, int) { return e.k; }
__device__ __forceinline__ uint32_t word(const E64& e, int w) { return w ? (uint32_t)(e.k >> 32) : (uint32_t)e.k; }
__device__ __forceinline__ uint32_t word(const P96& e, int w) {
  return w == 2 ? e.v : (w ? (uint32_t)(e.k >> 32) : (uint32_t)e.k);
}
template <class E> __device__ __forceinline__ E from_words(const uint32_t* w);
template <> __device__ __forceinline__ E32 from_words<E32>(const uint32_t* w) { return {w[0]}; }
template <> __device__ __forceinline__ E64 from_words<E64>(const uint32_t* w) {
  return {(uint64_t)w[0] | ((uint64_t)w[1] << 32)};
}
template <> __device__ __forceinline__ P96 from_words<P96>(const uint32_t* w) {
  return {(uint64_t)w[0] | ((uint64_t)w[1] << 32), w[2]};
}

template <int M, class E>
__device__ __forceinline__ E xlane(const E& e) {
  uint32_t w[E::W];
#pragma unroll
  for (int i = 0; i < E::W; ++i) w[i] = xor_lane<M>(word(e, i));
  return from_words<E>(w);
}

// ------------------------------------------------------------ dtype maps
template <int DT> struct DType;
template <> struct DType<DT_I32> {
  using E = E32; using Word = uint32_t;
  __device__ static uint32_t in(uint32_t x) { return x ^ 0x80000000u; }
  __device__ static uint32_t out(uint32_t u) { return u ^ 0x80000000u; }
};
template <> struct DType<DT_U32> {
  using E = E32; using Word = uint32_t;
  __device__ static uint32_t in(uint32_t x) { return x; }
  __device__ static uint32_t out(uint32_t u) { return u; }
};
template <> struct DType<DT_F32> {
  using E = E32; using Word = uint32_t;
  __device__ static uint32_t in(uint32_t x) { return x ^ ((uint32_t)((int32_t)x >> 31) | 0x80000000u); }
  __device__ static uint32_t out(uint32_t u) { return u ^ ((uint32_t)((int32_t)~u >> 31) | 0x80000000u); }
};
template <> struct DType<DT_I64> {
  using E = E64; using Word = uint64_t;
  __device__ static uint64_t in(uint64_t x) { return x ^ 0x8000000000000000ull; }
  __device__ static uint64_t out(uint64_t u) { return u ^ 0x8000000000000000ull; }
};
template <> struct DType<DT_U64> {
  using E = E64; using Word = uint64_t;
  __device__ static uint64_t in(uint64_t x) { return x; }
  __device__ static uint64_t out(uint64_t u) { return u; }
};
template <> struct DType<DT_F64> {
  using E = E64; using Word = uint64_t;
  __device__ static uint64_t in(uint64_t x) { return x ^ ((uint64_t)((int64_t)x >> 63) | 0x8000000000000000ull); }
  __device__ static uint64_t out(uint64_t u) { return u ^ ((uint64_t)((int64_t)~u >> 63) | 0x8000000000000000ull); }
};
template <> struct DType<DT_PAIR> {
  using E = P96; using Word = uint64_t;
  __device__ static uint64_t in(uint64_t x) { return x; }
  __device__ static uint64_t out(uint64_t u) { return u; }
};

// Global element access. keys: Word array; vals: u32 array (pairs only).
// Global load/store of one element; CONV maps the caller's bits to the
// internal order and back. flip (0 or all ones) XORs every internal word on
// that mapping: all ones reverses the order (descending sorts), so the
// kernels in between never know.
template <int DT, bool CONV>
__device__ __forceinline__ typename DType<DT>::E gload(const void* keys, const uint32_t* vals, int64_t i,
                                                      uint32_t flip = 0u) {
  using D = DType<DT>;
  using E = typename D::E;
  using Wd = typename D::Word;
  Wd x = static_cast<const Wd*>(keys)[i];
  if (CONV) x = D::in(x) ^ (Wd)(int32_t)flip;  // sign-extends to all 64 bits
  E e;
  if constexpr (E::HAS_VAL) {
    e.k = x;
    e.v = CONV ? vals[i] ^ flip : vals[i];
  } else {
    e.k = x;
  }
  return e;
}
template <int DT, bool CONV>
__device__ __forceinline__ void gstore(void* keys, uint32_t* vals, int64_t i, const typename DType<DT>::E& e,
                                       uint32_t flip = 0u) {
  using D = DType<DT>;
  using E = typename D::E;
  using Wd = typename D::Word;
  Wd x = (Wd)e.k;
  if (CONV) x = D::out(x ^ (Wd)(int32_t)flip);
  static_cast<Wd*>(keys)[i] = x;
  if constexpr (E::HAS_VAL) vals[i] = CONV ? e.v ^ flip : e.v;
}
// Four consecutive 32-bit elements base + o .. base + o + 3 into e[0..3]
// with one 16-byte load (element alignment suffices).
template <int DT, bool CONV>
__device__ __forceinline__ void gload4(const void* keys, int64_t base, int o, typename DType<DT>::E* e,
                                       uint32_t flip = 0u) {
  using D = DType<DT>;
  static_assert(D::E::W == 1, "32-bit keys");
  typedef uint32_t U4 __attribute__((ext_vector_type(4), aligned(4)));
  const U4 x = *(const U4*)(static_cast<const uint32_t*>(keys) + base + o);
#pragma unroll
  for (int t = 0; t < 4; ++t) e[t].k = CONV ? D::in(x[t]) ^ flip : x[t];
}
// Two consecutive elements of 64-bit keys (and their values) base + o,
// base + o + 1 into e[0..1] with one 16-byte key load (and one 8-byte value
// load); element alignment suffices.
template <int DT, bool CONV>
__device__ __forceinline__ void gload2(const void* keys, const uint32_t* vals, int64_t base, int o,
                                       typename DType<DT>::E* e, uint32_t flip = 0u) {
  using D = DType<DT>;
  using E = typename D::E;
  using Wd = typename D::Word;
  static_assert(sizeof(Wd) == 8, "64-bit keys");
  typedef Wd K2 __attribute__((ext_vector_type(2), aligned(8)));
  const K2 x = *(const K2*)(static_cast<const Wd*>(keys) + base + o);
#pragma unroll
  for (int t = 0; t < 2; ++t) e[t].k = CONV ? D::in(x[t]) ^ (Wd)(int32_t)flip : x[t];
  if constexpr (E::HAS_VAL) {
    typedef uint32_t V2 __attribute__((ext_vector_type(2), aligned(4)));
    const V2 y = *(const V2*)(vals + base + o);
#pragma unroll
    for (int t = 0; t < 2; ++t) e[t].v = CONV ? y[t] ^ flip : y[t];
  }
}
// Four consecutive elements e[0..3] to positions base + o .. base + o + 3
// with 16-byte vector stores (two for 64-bit keys). base is wave-uniform and
// the addresses need only element alignment.
template <int DT, bool CONV>
__device__ __forceinline__ void gstore4(void* keys, uint32_t* vals, int64_t base, int o,
                                        const typename DType<DT>::E* e, uint32_t flip = 0u) {
  using D = DType<DT>;
  using E = typename D::E;
  using Wd = typename D::Word;
  typedef Wd Wd4 __attribute__((ext_vector_type(4), aligned(sizeof(Wd))));
  Wd4 x;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    Wd k = (Wd)e[t].k;
    if (CONV) k = D::out(k ^ (Wd)(int32_t)flip);
    x[t] = k;
  }
  *(Wd4*)(static_cast<Wd*>(keys) + base + o) = x;
  if constexpr (E::HAS_VAL) {
    typedef uint32_t V4 __attribute__((ext_vector_type(4), aligned(4)));
    V4 y;
#pragma unroll
    for (int t = 0; t < 4; ++t) y[t] = CONV ? e[t].v ^ flip : e[t].v;
    *(V4*)(vals + base + o) = y;
  }
}

// Two consecutive 32-bit elements e[0..1] to base + o, base + o + 1 with one
// 8-byte store (element alignment suffices).
template <int DT, bool CONV>
__device__ __forceinline__ void gstore2(void* keys, int64_t base, int o, const typename DType<DT>::E* e,
                                        uint32_t flip = 0u) {
  using D = DType<DT>;
  static_assert(D::E::W == 1, "32-bit keys");
  using Wd = typename D::Word;
  typedef Wd Wd2 __attribute__((ext_vector_type(2), aligned(sizeof(Wd))));
  Wd2 x;
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    Wd k = (Wd)e[t].k;
    if (CONV) k = D::out(k ^ (Wd)(int32_t)flip);
    x[t] = k;
  }
  *(Wd2*)(static_cast<Wd*>(keys) + base + o) = x;
}

}  // namespace gmms
