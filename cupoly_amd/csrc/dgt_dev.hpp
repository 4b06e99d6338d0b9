// Device-side DGT machinery shared by the transform kernels (dgt.hip), the
// fused sample-and-multiply kernel (encrypt.hip) and the persistent uneven
// batch kernel (uneven.hip): register/LDS layouts of the radix-8 passes, the
// butterflies, the pass plan and the twist.  See dgt.hip for the design.
#pragma once
#include "internal.hpp"

#define EPT 8       // Gaussian slots per thread
// A scheduling barrier after every radix-2 stage of a pass keeps one stage's
// twiddle loads in flight at a time.  Since the passes are unrolled, the
// compiler's own schedule across stages is faster for the products and the
// seven-transform hom-mul kernel (1.0% and 2.0%); the relinearization kernel,
// which holds more values live, keeps the barriers (relin.hip sets 1).
#ifndef CP_STAGE_BARRIER
#define CP_STAGE_BARRIER 0
#endif
#define LOG_EPT 3

// -------------------------------------------------------------------------
// Index maps: local slot index (inside the workgroup's LDS array) -> global
// slot index inside one residue polynomial.
// -------------------------------------------------------------------------
// kAligned: global index == local index modulo 8 (true for whole polynomials
// and 128-slot blocks), so in a last pass of unit stride 1 the twiddles of the
// final three stages are known at compile time.
struct MapIdentity {  // whole polynomial, local == global
  static constexpr bool kAligned = true;
  static constexpr bool kLocalMask = true;
  __device__ __forceinline__ int operator()(int l) const { return l; }
};
template <int LOGC, int LOGNS>  // 2^LOGC columns of a polynomial, column stride 2^LOGNS
struct MapColumns {
  static constexpr bool kAligned = false;
  static constexpr bool kLocalMask = false;
  int col0;
  __device__ __forceinline__ int operator()(int l) const {
    return col0 + (l & ((1 << LOGC) - 1)) + ((l >> LOGC) << LOGNS);
  }
};
struct MapBlock {  // contiguous block starting at base (a multiple of the local size)
  static constexpr bool kAligned = true;
  static constexpr bool kLocalMask = true;
  int base;
  __device__ __forceinline__ int operator()(int l) const { return base + l; }
};

__device__ __forceinline__ void ld_tw(const uint64_t *__restrict__ tw, unsigned e, uint64_t &w,
                                      uint64_t &wp) {
  // 32-bit byte offset from a (usually uniform) base: saddr addressing
  const ulonglong2 v =
      *reinterpret_cast<const ulonglong2 *>(reinterpret_cast<const char *>(tw) + (e << 4));
  w = v.x;
  wp = v.y;
}
__device__ __forceinline__ cp_gi ld_gi(const CP_GLOBAL cp_gi *ptr) {
  const cp_u64x2 v = *(const CP_GLOBAL cp_u64x2 *)ptr;
  return cp_gi{v.x, v.y};
}
__device__ __forceinline__ void st_gi(CP_GLOBAL cp_gi *ptr, const cp_gi &v) {
  *(CP_GLOBAL cp_u64x2 *)ptr = cp_u64x2{v.re, v.imag};
}
__device__ __forceinline__ void ld_tw(const CP_GLOBAL uint64_t *__restrict__ tw, unsigned e,
                                      uint64_t &w, uint64_t &wp) {
  const cp_u64x2 v = *(const CP_GLOBAL cp_u64x2 *)(tw + 2 * e);
  w = v.x;
  wp = v.y;
}

// A twiddle table whose base is wave-uniform, also held as a buffer
// resource: a load whose index is a per-lane part plus a compile-time part
// takes the first as a 32-bit VGPR offset and the second as an SGPR offset
// (scalar adds), with no 64-bit address arithmetic per load.  Passed to the
// pass functions by pointer (a local of the caller; it lives in SGPRs once
// inlined).  Out-of-range buffer reads return 0: `bytes` is the table size.
struct UTab {
  __amdgpu_buffer_rsrc_t r;
  const uint64_t *p;
};
__device__ __forceinline__ UTab utab(const uint64_t *p, unsigned bytes) {
  // raw buffer, no stride: the gfx9 resource word 3
  return UTab{__builtin_amdgcn_make_buffer_rsrc((void *)p, (short)0, (int)bytes, 0x00020000), p};
}
// uniform index: the pointer path (a scalar load when the index is uniform)
__device__ __forceinline__ void ld_tw(const UTab *tw, unsigned e, uint64_t &w, uint64_t &wp) {
  ld_tw(tw->p, e, w, wp);
}
// index e_lane + e_const, e_const known at compile time
__device__ __forceinline__ void ld_tw_lc(const UTab *tw, unsigned e_lane, unsigned e_const,
                                         uint64_t &w, uint64_t &wp) {
  const cp_u64x2 v = __builtin_bit_cast(
      cp_u64x2, __builtin_amdgcn_raw_buffer_load_b128(tw->r, (int)(e_lane << 4),
                                                      (int)(e_const << 4), 0));
  w = v.x;
  wp = v.y;
}
template <class P>
__device__ __forceinline__ void ld_tw_lc(P tw, unsigned e_lane, unsigned e_const, uint64_t &w,
                                         uint64_t &wp) {
  ld_tw(tw, e_lane + e_const, w, wp);
}

// Wave-uniform pointer: kept in SGPRs (the compiler otherwise holds some
// uniform bases in VGPR pairs across the transforms, where they spill).  The
// rebuilt address is first typed as a global-memory pointer, so the accesses
// through it stay global loads and stores: a plain integer-to-pointer cast
// gives a generic pointer, whose flat accesses also count against the LDS
// wait counter and so stall the exchanges' LDS waits.
template <class P>
__device__ __forceinline__ P *uniform_ptr(P *q) {
  const uint64_t v = (uint64_t)q;
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  typedef __attribute__((address_space(1))) P global_p;
  return (P *)(global_p *)(((uint64_t)hi << 32) | lo);
}

// One forward pass over local stages [LS0, LS0+R) of a 2^LOGM local array.
// Global stage = local stage + SOFF; LOGN = log2 of the global transform size.
// Butterflies whose twiddle is 1 are done without a multiply: always in the
// last global stage, and, when the unit stride is 1 on an aligned map, every
// butterfly whose index within its group is 0.
template <int LOGN, int LOGM, int LS0, int R, int SOFF, class Map, class TW>
__device__ __forceinline__ void dif_pass(cp_gi (&x)[EPT], int t, TW __restrict__ tw,
                                         uint64_t p, const Map &map) {
  constexpr int M = 1 << LOGM, T = M / EPT, U = 1 << R, ST = M >> (LS0 + R), NU = EPT / U;
#pragma unroll
  for (int ju = 0; ju < NU; ju++) {
    const int u = t + ju * T;
    const int base = (u / ST) * (U * ST) + (u % ST);
#pragma unroll
    for (int h = 0; h < R; h++) {
      const int half = U >> (h + 1);
      const int s = LS0 + h + SOFF;
      const int m = 1 << (LOGN - 1 - s);
#pragma unroll
      for (int q = 0; q < U; q++) {
        if (q & half) continue;
        cp_gi &A = x[ju * U + q];
        cp_gi &B = x[ju * U + q + half];
        const uint64_t sr = A.re + B.re, si = A.imag + B.imag;
        const uint64_t dr = A.re + p - B.re, di = A.imag + p - B.imag;
        A.re = cp_red2(sr, p);
        A.imag = cp_red2(si, p);
        const bool static_j = Map::kAligned && ST == 1 && m <= U;
        const int jc = q & (m - 1);
        if (s == LOGN - 1 || (static_j && jc == 0)) {  // twiddle 1
          B.re = cp_red2(dr, p);
          B.imag = cp_red2(di, p);
        } else {
          const int e = static_j ? (jc << s) : ((map(base + q * ST) & (m - 1)) << s);
          uint64_t w, wp;
          if (static_j) {  // a uniform twiddle (may sit in SGPRs)
            ld_tw(tw, e, w, wp);
            cp_shoup2(dr, w, wp, di, w, wp, p, B.re, B.imag);
          } else {
            ld_tw_lc(tw, e, 0u, w, wp);
            cp_shoup2v(dr, w, wp, di, w, wp, p, B.re, B.imag);
          }
        }
      }
      if (CP_STAGE_BARRIER) __builtin_amdgcn_sched_barrier(0);
    }
  }
}

// One inverse pass: same units, stages in reverse, DIT butterflies with omega^-1.
template <int LOGN, int LOGM, int LS0, int R, int SOFF, class Map, class TW>
__device__ __forceinline__ void dit_pass(cp_gi (&x)[EPT], int t, TW __restrict__ tw,
                                         uint64_t p, const Map &map) {
  constexpr int M = 1 << LOGM, T = M / EPT, U = 1 << R, ST = M >> (LS0 + R), NU = EPT / U;
#pragma unroll
  for (int ju = 0; ju < NU; ju++) {
    const int u = t + ju * T;
    const int base = (u / ST) * (U * ST) + (u % ST);
#pragma unroll
    for (int h = R - 1; h >= 0; h--) {
      const int half = U >> (h + 1);
      const int s = LS0 + h + SOFF;
      const int m = 1 << (LOGN - 1 - s);
#pragma unroll
      for (int q = 0; q < U; q++) {
        if (q & half) continue;
        cp_gi &A = x[ju * U + q];
        cp_gi &B = x[ju * U + q + half];
        uint64_t tr, ti;
        const bool static_j = Map::kAligned && ST == 1 && m <= U;
        const int jc = q & (m - 1);
        if (s == LOGN - 1 || (static_j && jc == 0)) {
          tr = B.re;
          ti = B.imag;
        } else {
          const int e = static_j ? (jc << s) : ((map(base + q * ST) & (m - 1)) << s);
          uint64_t w, wp;
          ld_tw(tw, e, w, wp);
          if (static_j)
            cp_shoup2(B.re, w, wp, B.imag, w, wp, p, tr, ti);
          else
            cp_shoup2v(B.re, w, wp, B.imag, w, wp, p, tr, ti);
        }
        B.re = cp_red2(A.re + p - tr, p);
        B.imag = cp_red2(A.imag + p - ti, p);
        A.re = cp_red2(A.re + tr, p);
        A.imag = cp_red2(A.imag + ti, p);
      }
      if (CP_STAGE_BARRIER) __builtin_amdgcn_sched_barrier(0);
    }
  }
}

// Local index of register slot i in the layout of the pass starting at LS0 with R stages.
template <int LOGM, int LS0, int R>
__device__ __forceinline__ int pass_index(int t, int i) {
  constexpr int M = 1 << LOGM, T = M / EPT, U = 1 << R, ST = M >> (LS0 + R);
  const int ju = i / U, q = i % U;
  const int u = t + ju * T;
  return (u / ST) * (U * ST) + (u % ST) + q * ST;
}

// Same for a full radix-8 pass (R = 3, one unit per thread) starting at a
// runtime local stage ls0: unit stride ST = M >> (ls0 + 3).
template <int LOGM>
__device__ __forceinline__ int pass_index_rt(int t, int i, int ls0) {
  const int lst = LOGM - ls0 - LOG_EPT;  // log2(ST)
  return ((t >> lst) << (lst + LOG_EPT)) + (t & ((1 << lst) - 1)) + (i << lst);
}

// ---- LDS layout of an exchange ---------------------------------------------
// An exchange writes the registers in one pass layout and reads them in the
// next.  A 16-lane group of 16-byte accesses is conflict-free when it hits 16
// distinct slots modulo 256 bytes; layouts with unit stride ST = 1 (8
// consecutive slots per lane) or ST = 8 (runs of 8 lanes 64 slots apart) hit
// only 2 of them.  Each exchange therefore pads the array linearly,
//     physical = l + A * (l >> S),
// with (A, S) = (1, 3) when the smaller stride of its two layouts is 1, and
// (8, 6) when it is 8, 64, 512 or 4096.  (8, 6) is needed at stride 8; at
// the larger strides 16 lanes touch 16 consecutive slots inside one 64-slot
// run, which it leaves consecutive, so it costs nothing there and lets the
// exchanges of a pass sequence share one map (below: the barrier before the
// stores can then go).  Other strides (16, 32) take no padding.  The maps
// are conflict-free for both layouts of their exchange (and for the natural
// order used to stage global I/O), and keep every address a per-thread base
// plus i times a uniform stride (S <= the stride's log2, or S = it + 3).  A
// 2^LOGM-slot array then spans 2^LOGM + 2^LOGM / 8 slots.
struct LdsMap {
  int A, S;
};
__host__ __device__ constexpr LdsMap lds_map_for(int lst_lo) {
  return lst_lo == 0 ? LdsMap{1, 3} : (lst_lo >= 3 && lst_lo % 3 == 0) ? LdsMap{8, 6} : LdsMap{0, 0};
}
template <int LOGM>
__host__ __device__ constexpr int lds_row() {
  return (1 << LOGM) + ((1 << LOGM) >> 3);
}
__device__ __forceinline__ int lds_phys(int l, LdsMap m) { return l + m.A * (l >> m.S); }

template <int LOGM, int LS0, int R, int A, int S>
__device__ __forceinline__ void lds_store(cp_gi *lds, const cp_gi (&x)[EPT], int t) {
#pragma unroll
  for (int i = 0; i < EPT; i++) {
    const int l = pass_index<LOGM, LS0, R>(t, i);
    lds[l + A * (l >> S)] = x[i];
  }
}
template <int LOGM, int LS0, int R, int A, int S>
__device__ __forceinline__ void lds_load(const cp_gi *lds, cp_gi (&x)[EPT], int t) {
#pragma unroll
  for (int i = 0; i < EPT; i++) {
    const int l = pass_index<LOGM, LS0, R>(t, i);
    x[i] = lds[l + A * (l >> S)];
  }
}
// Full radix-8 layout at runtime stage ls0 (stride 2^lst) under map m, which
// must have S <= lst or S == lst + 3 (true for the maps chosen above):
// physical(base + i 2^lst) = PB + i D.
template <int LOGM>
__device__ __forceinline__ void lds_rt_addr(int t, int ls0, LdsMap m, int &pb, int &d) {
  const int lst = LOGM - ls0 - LOG_EPT;
  const int b = ((t >> lst) << (lst + LOG_EPT)) + (t & ((1 << lst) - 1));
  pb = b + m.A * (b >> m.S);
  d = (1 << lst) + (lst >= m.S ? m.A << (lst - m.S) : 0);
}
template <int LOGM>
__device__ __forceinline__ void lds_store_rt(cp_gi *lds, const cp_gi (&x)[EPT], int t, int ls0,
                                             LdsMap m) {
  int pb, d;
  lds_rt_addr<LOGM>(t, ls0, m, pb, d);
#pragma unroll
  for (int i = 0; i < EPT; i++) lds[pb + i * d] = x[i];
}
template <int LOGM>
__device__ __forceinline__ void lds_load_rt(const cp_gi *lds, cp_gi (&x)[EPT], int t, int ls0,
                                            LdsMap m) {
  int pb, d;
  lds_rt_addr<LOGM>(t, ls0, m, pb, d);
#pragma unroll
  for (int i = 0; i < EPT; i++) x[i] = lds[pb + i * d];
}

// Full radix-8 pass at a runtime local stage ls0 (never the last global
// stage, so every twiddle is a real multiply).
template <int LOGN, int LOGM, int SOFF, class Map, class TW>
__device__ __forceinline__ void dif_pass_rt(cp_gi (&x)[EPT], int t, TW __restrict__ tw,
                                            uint64_t p, const Map &map, int ls0) {
  const int lst = LOGM - ls0 - LOG_EPT;
  const int base = ((t >> lst) << (lst + LOG_EPT)) + (t & ((1 << lst) - 1));
  const unsigned t0 = (unsigned)t & ((1u << lst) - 1);
#pragma unroll
  for (int h = 0; h < LOG_EPT; h++) {
    const int half = EPT >> (h + 1);
    const int s = ls0 + h + SOFF;
    const int mmask = (1 << (LOGN - 1 - s)) - 1;
    const unsigned e0 = t0 << s;
#pragma unroll
    for (int q = 0; q < EPT; q++) {
      if (q & half) continue;
      cp_gi &A = x[q];
      cp_gi &B = x[q + half];
      const uint64_t sr = A.re + B.re, si = A.imag + B.imag;
      const uint64_t dr = A.re + p - B.re, di = A.imag + p - B.imag;
      A.re = cp_red2(sr, p);
      A.imag = cp_red2(si, p);
      uint64_t w, wp;
      if (Map::kLocalMask)
        ld_tw_lc(tw, e0, (unsigned)(q & (half - 1)) * (1u << (LOGM - LOG_EPT + SOFF + h)), w, wp);
      else
        ld_tw(tw, (unsigned)((map(base + (q << lst)) & mmask) << s), w, wp);
      cp_shoup2v(dr, w, wp, di, w, wp, p, B.re, B.imag);
    }
    // one stage's twiddles in flight at a time (bounds VGPR use)
    if (CP_STAGE_BARRIER) __builtin_amdgcn_sched_barrier(0);
  }
}

template <int LOGN, int LOGM, int SOFF, class Map, class TW>
__device__ __forceinline__ void dit_pass_rt(cp_gi (&x)[EPT], int t, TW __restrict__ tw,
                                            uint64_t p, const Map &map, int ls0) {
  const int lst = LOGM - ls0 - LOG_EPT;
  const int base = ((t >> lst) << (lst + LOG_EPT)) + (t & ((1 << lst) - 1));
  const unsigned t0 = (unsigned)t & ((1u << lst) - 1);
#pragma unroll
  for (int h = LOG_EPT - 1; h >= 0; h--) {
    const int half = EPT >> (h + 1);
    const int s = ls0 + h + SOFF;
    const int mmask = (1 << (LOGN - 1 - s)) - 1;
    const unsigned e0 = t0 << s;
#pragma unroll
    for (int q = 0; q < EPT; q++) {
      if (q & half) continue;
      cp_gi &A = x[q];
      cp_gi &B = x[q + half];
      const unsigned e = Map::kLocalMask
                             ? e0 + (unsigned)(q & (half - 1)) * (1u << (LOGM - LOG_EPT + SOFF + h))
                             : (unsigned)((map(base + (q << lst)) & mmask) << s);
      uint64_t w, wp;
      ld_tw(tw, e, w, wp);
      uint64_t tr, ti;
      cp_shoup2v(B.re, w, wp, B.imag, w, wp, p, tr, ti);
      B.re = cp_red2(A.re + p - tr, p);
      B.imag = cp_red2(A.imag + p - ti, p);
      A.re = cp_red2(A.re + tr, p);
      A.imag = cp_red2(A.imag + ti, p);
    }
    if (CP_STAGE_BARRIER) __builtin_amdgcn_sched_barrier(0);
  }
}

constexpr int cmin(int a, int b) { return a < b ? a : b; }

// An opaque copy of a thread index: address arithmetic derived from it
// cannot be hoisted above this point, nor shared between two pass sequences
// and held in VGPRs across the transform in between (where it would spill).
__device__ __forceinline__ int cp_opaque(int t) {
  asm volatile("" : "+v"(t));
  return t;
}

// ---- Exchange synchronisation ------------------------------------------------
// An exchange between full radix-8 layouts at local stages ls0 and ls0 + 3
// is wave-local when the earlier layout's unit stride 2^lst is at most 64
// (lst = LOGM - ls0 - 3 <= 6): each group of 2^lst threads then owns one
// block of 8 * 2^lst consecutive slots, so the 64 threads of a wave own the
// same 512 slots (or the same whole rows, when a row has 64 threads or fewer)
// in both layouts, and read back only what their own wave wrote.  Such an
// exchange needs no workgroup barrier between its stores and its loads: LDS
// instructions of one wave execute in program order, and wave_sync() keeps
// the compiler from moving them across.  The barrier before its stores (no
// wave may still be reading the region it overwrites) can also go when the
// previous exchange was wave-local under a map with the same padding ratio
// (1/8 or none), because every wave then read, and now writes, only its own
// region.  Everywhere else both barriers stay.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
__device__ __forceinline__ bool wave_local(int lst) { return lst <= 6; }
// The barrier before an exchange's stores can also go when the previous
// exchange of the same pass sequence used the same LDS map: each thread then
// stores its registers, in the layout it loaded them in, to exactly the LDS
// words it read itself, so no other wave can still be reading a word it
// overwrites (and one wave's LDS accesses run in program order).  With the
// maps above, every exchange of an N = 8192-slot sequence but the last
// (stride 1) uses (8, 6).
__device__ __forceinline__ bool same_map(LdsMap a, LdsMap b) { return a.A == b.A && a.S == b.S; }

// Pass plan over local stages [LSB, LSE): an optional partial pass of
// R0 = NS mod 3 stages first, then NF full radix-8 passes.  The full passes
// but the last run in a loop that is not unrolled (so the compiler cannot
// hoist twiddle loads across passes and spill); the last pass is compiled
// with its stages known, which lets its trivial twiddles drop out.
template <int LSB, int LSE>
struct PassPlan {
  static constexpr int NS = LSE - LSB;
  static constexpr int R0 = NS % LOG_EPT;                  // partial first pass (0: none)
  static constexpr int NF = NS / LOG_EPT;                  // full passes
  static constexpr int LSF = LSB + R0;                     // first stage of the full passes
  static constexpr int LSL = NF ? LSE - LOG_EPT : LSB;     // first stage of the last pass
  static constexpr int RL = NF ? LOG_EPT : R0;             // stages in the last pass
  static constexpr int LS_FIRST = LSB;                     // first pass
  static constexpr int R_FIRST = R0 ? R0 : LOG_EPT;
};

// Register layout of the first pass (used for direct global I/O).
template <int LOGM, int LSB, int LSE>
__device__ __forceinline__ int first_index(int t, int i) {
  using PP = PassPlan<LSB, LSE>;
  return pass_index<LOGM, PP::LS_FIRST, PP::R_FIRST>(t, i);
}

// Called by a pass sequence after its last LDS read, before its last pass
// (which runs in registers): the LDS array is free from there on for the
// workgroup's waves once they have all passed a barrier.
struct NoHook {
  __device__ __forceinline__ void operator()() const {}
};

// Forward passes over local stages [LSB, LSE), registers starting in the
// first-pass layout and ending in the last-pass layout.
template <int LOGN, int LOGM, int LSB, int LSE, int SOFF, class Map, class TW,
          class Hook = NoHook>
__device__ __forceinline__ void dif_passes(cp_gi (&x)[EPT], cp_gi *lds, int t,
                                           TW __restrict__ tw, uint64_t p,
                                           const Map &map, const Hook &hook = Hook()) {
  using PP = PassPlan<LSB, LSE>;
  constexpr LdsMap MP = lds_map_for(LOGM - PP::LSF - LOG_EPT);  // partial <-> first full
  t = cp_opaque(t);
  LdsMap prev_m{-1, -1};  // map of the previous exchange of this sequence (none yet)
  if constexpr (PP::R0 > 0) {
    dif_pass<LOGN, LOGM, LSB, PP::R0, SOFF>(x, t, tw, p, map);
    if constexpr (PP::NF == 0) return;
    __syncthreads();
    lds_store<LOGM, LSB, PP::R0, MP.A, MP.S>(lds, x, t);
    __syncthreads();
    lds_load_rt<LOGM>(lds, x, t, PP::LSF, MP);
    prev_m = MP;
  }
  bool prev_local = false, prev_pad = false;  // previous exchange of this loop
  // unrolled: each pass's stage, twiddle indices, LDS map and barrier kind
  // are then compile-time (immediate LDS offsets, no index arithmetic per
  // slot): 4.6% faster products than the rolled loop, at 1.6x the code
#pragma unroll
  for (int i = 0; i < PP::NF - 1; i++) {
    const int ls0 = PP::LSF + LOG_EPT * i;
    const LdsMap m = lds_map_for(LOGM - ls0 - 2 * LOG_EPT);
    const bool local = wave_local(LOGM - ls0 - LOG_EPT), pad = m.A != 0;
    dif_pass_rt<LOGN, LOGM, SOFF>(x, t, tw, p, map, ls0);
    if ((local && prev_local && pad == prev_pad) || same_map(m, prev_m)) wave_sync(); else __syncthreads();
    lds_store_rt<LOGM>(lds, x, t, ls0, m);
    if (local) wave_sync(); else __syncthreads();
    lds_load_rt<LOGM>(lds, x, t, ls0 + LOG_EPT, m);
    prev_local = local;
    prev_pad = pad;
    prev_m = m;
  }
  hook();
  if constexpr (PP::NF > 0) dif_pass<LOGN, LOGM, PP::LSL, LOG_EPT, SOFF>(x, t, tw, p, map);
}

// Inverse passes over local stages [LSB, LSE) in reverse order; registers
// start in the last-pass layout and end in the first-pass layout.
template <int LOGN, int LOGM, int LSB, int LSE, int SOFF, class Map, class TW>
__device__ __forceinline__ void dit_passes(cp_gi (&x)[EPT], cp_gi *lds, int t,
                                           TW __restrict__ tw, uint64_t p,
                                           const Map &map) {
  using PP = PassPlan<LSB, LSE>;
  constexpr LdsMap MP = lds_map_for(LOGM - PP::LSF - LOG_EPT);
  t = cp_opaque(t);
  if constexpr (PP::NF > 0) dit_pass<LOGN, LOGM, PP::LSL, LOG_EPT, SOFF>(x, t, tw, p, map);
  bool prev_local = false, prev_pad = false;
  LdsMap prev_m{-1, -1};  // map of the previous exchange of this sequence (none yet)
#pragma unroll  // as in dif_passes
  for (int i = PP::NF - 2; i >= 0; i--) {
    const int ls0 = PP::LSF + LOG_EPT * i;
    const LdsMap m = lds_map_for(LOGM - ls0 - 2 * LOG_EPT);
    const bool local = wave_local(LOGM - ls0 - LOG_EPT), pad = m.A != 0;
    if ((local && prev_local && pad == prev_pad) || same_map(m, prev_m)) wave_sync(); else __syncthreads();
    lds_store_rt<LOGM>(lds, x, t, ls0 + LOG_EPT, m);
    if (local) wave_sync(); else __syncthreads();
    lds_load_rt<LOGM>(lds, x, t, ls0, m);
    dit_pass_rt<LOGN, LOGM, SOFF>(x, t, tw, p, map, ls0);
    prev_local = local;
    prev_pad = pad;
    prev_m = m;
  }
  if constexpr (PP::R0 > 0) {
    if constexpr (PP::NF > 0) {
      if (same_map(MP, prev_m)) wave_sync(); else __syncthreads();
      lds_store_rt<LOGM>(lds, x, t, PP::LSF, MP);
      __syncthreads();
      lds_load<LOGM, LSB, PP::R0, MP.A, MP.S>(lds, x, t);
    }
    dit_pass<LOGN, LOGM, LSB, PP::R0, SOFF>(x, t, tw, p, map);
  }
}

// Layout of the last pass over [LSB, LSE).
template <int LOGM, int LSB, int LSE>
struct LastPass {
  static constexpr int LS0 = PassPlan<LSB, LSE>::LSL;
  static constexpr int RL = PassPlan<LSB, LSE>::RL;
  // map of an exchange between this layout and the natural order
  static constexpr LdsMap MS = lds_map_for(LOGM - LS0 - RL);
};

// natural layout index of register i: t + i*T (LDS-staged global I/O)
template <int LOGM>
__device__ __forceinline__ int nat_index(int t, int i) {
  return t + i * ((1 << LOGM) / EPT);
}

__device__ __forceinline__ void twist(cp_gi &v, const uint64_t *__restrict__ tab, int g, uint64_t p) {
  const ulonglong2 *k2 = reinterpret_cast<const ulonglong2 *>(tab + 6 * (size_t)g);
  const ulonglong2 a = k2[0], b = k2[1], c = k2[2];
  const uint64_t k[6] = {a.x, a.y, b.x, b.y, c.x, c.y};
  v = cp_gimul_const(v, k, p);
}
__device__ __forceinline__ void twist(cp_gi &v, const CP_GLOBAL uint64_t *__restrict__ tab, int g,
                                      uint64_t p) {
  const CP_GLOBAL cp_u64x2 *k2 = (const CP_GLOBAL cp_u64x2 *)(tab + 6 * (size_t)g);
  const cp_u64x2 a = k2[0], b = k2[1], c = k2[2];
  const uint64_t k[6] = {a.x, a.y, b.x, b.y, c.x, c.y};
  v = cp_gimul_const(v, k, p);
}


// ---- Split-basis products --------------------------------------------------
// omega of order N >= 16 lies in Z_p, so p = 1 (mod 4): -1 has the square
// roots +-J in Z_p and Z_p[i] = Z_p x Z_p through a + i b -> (a + J b, a - J b).
// A product whose inputs and outputs are coefficients never exposes the DGT
// domain, so it may run in that basis.  Slot g = (a_g, a_{g+N}) becomes
// (u_g, v_g) = (a_g + J a_{g+N}, a_g - J a_{g+N}): the residues of a mod
// X^N - J and mod X^N + J.  Then
//   * the twist by xi^g (xi^N = i) is two real products, by xi+^g and xi-^g
//     (xi+- = xi.re +- J xi.im, so xi+^N = J and xi-^N = -J), in place of a
//     Gaussian product;
//   * the forward network (dif_passes, the same real omega) is unchanged and
//     leaves u(xi+ omega^k) and v(xi- omega^k) at position bitrev(k);
//   * the slot-wise product is two products mod p instead of a Gaussian one;
//   * the inverse undoes the split tree of X^N -+ J with Gentleman-Sande
//     butterflies, A + B and (A - B) r^-1, r the root that split the
//     butterfly's group (X^2m - r^2 = (X^m - r)(X^m + r)).  That absorbs the
//     untwist, and the difference feeds a Shoup product unreduced, where the
//     inverse DGT's butterflies reduce three times;
//   * back to coefficients, c_g = (U + V) / 2N and c_{g+N} = (U - V) / 2JN,
//     two Shoup products that also take their inputs unreduced.
// Tables per residue: twist [N][4] = (xi+^g, Shoup, xi-^g, Shoup); inverse
// roots [N][4] at node 2^s + group (stage s) = (r_u^-1, Shoup, r_v^-1, Shoup);
// constants [8] = (J, 1/2N, 1/2JN) as Shoup pairs.

// Forward conversion and twist of slots x (index gx) and y (index gy).
__device__ __forceinline__ void split_twist2(cp_gi &x, cp_gi &y, const CP_GLOBAL uint64_t *tab,
                                             int gx, int gy, uint64_t J, uint64_t Jp,
                                             uint64_t p) {
  uint64_t tx, ty;
  cp_shoup2(x.imag, J, Jp, y.imag, J, Jp, p, tx, ty);  // J may sit in SGPRs
  const CP_GLOBAL cp_u64x2 *kx = (const CP_GLOBAL cp_u64x2 *)(tab + 4 * (size_t)gx);
  const CP_GLOBAL cp_u64x2 *ky = (const CP_GLOBAL cp_u64x2 *)(tab + 4 * (size_t)gy);
  const cp_u64x2 px = kx[0], mx = kx[1], py = ky[0], my = ky[1];
  // a + t and a + p - t are below 2p: Shoup inputs as they are
  cp_shoup2v(x.re + tx, px.x, px.y, x.re + p - tx, mx.x, mx.y, p, x.re, x.imag);
  cp_shoup2v(y.re + ty, py.x, py.y, y.re + p - ty, my.x, my.y, p, y.re, y.imag);
}
// Same for one slot.
__device__ __forceinline__ void split_twist1(cp_gi &x, const CP_GLOBAL uint64_t *tab, int g,
                                             uint64_t J, uint64_t Jp, uint64_t p) {
  const uint64_t t = cp_shoup(x.imag, J, Jp, p);
  const CP_GLOBAL cp_u64x2 *k = (const CP_GLOBAL cp_u64x2 *)(tab + 4 * (size_t)g);
  const cp_u64x2 kp = k[0], km = k[1];
  cp_shoup2v(x.re + t, kp.x, kp.y, x.re + p - t, km.x, km.y, p, x.re, x.imag);
}
// Slot-wise product in the split basis.  split_mul_paired forms the two
// products with cp_mulmod2 (fewer instructions, more registers live at once:
// used where the registers allow it).
__device__ __forceinline__ cp_gi split_mul(const cp_gi &a, const cp_gi &b, const cp_prime &pr) {
  return cp_gi{cp_mulmod(a.re, b.re, pr), cp_mulmod(a.imag, b.imag, pr)};
}
__device__ __forceinline__ cp_gi split_mul_paired(const cp_gi &a, const cp_gi &b,
                                                  const cp_prime &pr) {
  cp_gi r;
  cp_mulmod2(a.re, b.re, a.imag, b.imag, pr, r.re, r.imag);
  return r;
}
// (U, V) -> (c_g, c_{g+N}) = ((U + V) h, (U - V) hj), h = 1/2N, hj = 1/2JN.
__device__ __forceinline__ void split_back(cp_gi &x, uint64_t h, uint64_t hp, uint64_t hj,
                                           uint64_t hjp, uint64_t p) {
  cp_shoup2(x.re + x.imag, h, hp, x.re + p - x.imag, hj, hjp, p, x.re, x.imag);
}
// The same for a product whose second operand was twisted with the table
// that carries h (d_split_tw_h): (U, V) already hold the factor h, so
// c_g = U + V and c_{g+N} = (U - V) J^-1 = (V - U) J (J^2 = -1).  Two slots,
// so that their products by J pair up.
__device__ __forceinline__ void split_back_h2(cp_gi &x, cp_gi &y, uint64_t J, uint64_t Jp,
                                              uint64_t p) {
  const uint64_t xr = x.re, yr = y.re;
  x.re = cp_red2(xr + x.imag, p);
  y.re = cp_red2(yr + y.imag, p);
  cp_shoup2(x.imag + p - xr, J, Jp, y.imag + p - yr, J, Jp, p, x.imag, y.imag);
}

__device__ __forceinline__ void ld_root2(const CP_GLOBAL uint64_t *__restrict__ tg, unsigned node,
                                         uint64_t &wu, uint64_t &wup, uint64_t &wv,
                                         uint64_t &wvp) {
  const CP_GLOBAL cp_u64x2 *k = (const CP_GLOBAL cp_u64x2 *)(tg + 4 * node);
  const cp_u64x2 a = k[0], b = k[1];
  wu = a.x;
  wup = a.y;
  wv = b.x;
  wvp = b.y;
}
__device__ __forceinline__ void ld_root2(const uint64_t *__restrict__ tg, unsigned node,
                                         uint64_t &wu, uint64_t &wup, uint64_t &wv,
                                         uint64_t &wvp) {
  const ulonglong2 *k = reinterpret_cast<const ulonglong2 *>(tg + 4 * node);
  const ulonglong2 a = k[0], b = k[1];
  wu = a.x;
  wup = a.y;
  wv = b.x;
  wvp = b.y;
}
// The root of a wave-uniform node (node 1 at stage 0, the same for every
// slot; any node whose group does not depend on the slot) through the
// constant address space: with a wave-uniform table pointer that is a scalar
// load, which stays off the vector memory counter (loads and LDS copies in
// flight are waited on through that counter, in order).  The tables do not
// change during a kernel.
template <class TW>
__device__ __forceinline__ void ld_root2_u(TW tg, unsigned node, uint64_t &wu, uint64_t &wup,
                                           uint64_t &wv, uint64_t &wvp) {
  typedef __attribute__((address_space(4))) const cp_u64x2 const_u64x2;
  const const_u64x2 *k = (const const_u64x2 *)(uintptr_t)(tg + 4 * node);
  const cp_u64x2 a = k[0], b = k[1];
  wu = a.x;
  wup = a.y;
  wv = b.x;
  wvp = b.y;
}

// One split-basis Gentleman-Sande butterfly: the group of A at global stage s
// is pos >> (LOGN - s), its root at node 2^s + group.  unode >= 0: the node,
// the same for every butterfly of the stage (a uniform root, possibly in
// SGPRs, which the compiler cannot see when pos comes from an opaque thread
// index).
template <int LOGN, class TW>
__device__ __forceinline__ void gsi_bfly(cp_gi &A, cp_gi &B, int s, int pos, TW __restrict__ tg,
                                         uint64_t p, int unode) {
  uint64_t wu, wup, wv, wvp;
  if (unode >= 0)
    ld_root2_u(tg, (unsigned)unode, wu, wup, wv, wvp);
  else
    ld_root2(tg, (1u << s) + ((unsigned)pos >> (LOGN - s)), wu, wup, wv, wvp);
  const uint64_t dr = A.re + p - B.re, di = A.imag + p - B.imag;
  A.re = cp_red2(A.re + B.re, p);
  A.imag = cp_red2(A.imag + B.imag, p);
  if (unode >= 0)
    cp_shoup2(dr, wu, wup, di, wv, wvp, p, B.re, B.imag);
  else
    cp_shoup2v(dr, wu, wup, di, wv, wvp, p, B.re, B.imag);
}

// One split-basis inverse pass over local stages [LS0, LS0+R) in reverse
// (the positions dit_pass pairs, in the same order).
template <int LOGN, int LOGM, int LS0, int R, int SOFF, class Map, class TW>
__device__ __forceinline__ void gsi_pass(cp_gi (&x)[EPT], int t, TW __restrict__ tg, uint64_t p,
                                         const Map &map) {
  constexpr int M = 1 << LOGM, T = M / EPT, U = 1 << R, ST = M >> (LS0 + R), NU = EPT / U;
#pragma unroll
  for (int ju = 0; ju < NU; ju++) {
    const int u = t + ju * T;
    const int base = (u / ST) * (U * ST) + (u % ST);
#pragma unroll
    for (int h = R - 1; h >= 0; h--) {
      const int half = U >> (h + 1);
      const int s = LS0 + h + SOFF;
#pragma unroll
      for (int q = 0; q < U; q++) {
        if (q & half) continue;
        // stage 0 has the single node 1 (pos < 2^LOGN); with a block map the
        // group of every slot is map(0) >> (LOGN - s) once LOGN - s >= LOGM
        const int unode = s == 0 ? 1
                          : (Map::kLocalMask && LOGN - s >= LOGM)
                              ? (int)((1u << s) + ((unsigned)map(0) >> (LOGN - s)))
                              : -1;
        gsi_bfly<LOGN>(x[ju * U + q], x[ju * U + q + half], s, map(base + q * ST), tg, p, unode);
      }
      if (CP_STAGE_BARRIER) __builtin_amdgcn_sched_barrier(0);
    }
  }
}

// Full radix-8 split-basis inverse pass at a runtime local stage ls0.
// UROOT and uroot: load wave-uniform roots as scalars where the nodes allow
// it (below); without UROOT the pass compiles as if that option did not exist.
template <int LOGN, int LOGM, int SOFF, bool UROOT = false, class Map, class TW>
__device__ __forceinline__ void gsi_pass_rt(cp_gi (&x)[EPT], int t, TW __restrict__ tg, uint64_t p,
                                            const Map &map, int ls0, bool uroot = false) {
  const int lst = LOGM - ls0 - LOG_EPT;
  const int base = ((t >> lst) << (lst + LOG_EPT)) + (t & ((1 << lst) - 1));
#pragma unroll
  for (int h = LOG_EPT - 1; h >= 0; h--) {
    const int half = EPT >> (h + 1);
    const int s = ls0 + h + SOFF;
    // block maps (map(l) = map(0) + l, map(0) a multiple of the local size,
    // LOGN = LOGM + SOFF): the group of local slot base + (q << lst) at this
    // stage is (map(0) >> (lst + 3 - h)) + ((t >> lst) << h) + (q >> (3 - h)),
    // since base has no bits in [lst, lst + 3): one node per stage plus a
    // compile-time offset per butterfly
    const unsigned node0 =
        (1u << s) + ((unsigned)map(0) >> (LOGN - s)) + (((unsigned)t >> lst) << h);
    // lst >= 6: t >> lst is the same for the 64 lanes of a wave, and so is
    // every node of the stage: with uroot its roots are scalar loads
    // (ld_root2_u), which leave the vector memory counter to loads a hook has
    // put in flight
    constexpr bool kBlock = Map::kLocalMask && LOGN == LOGM + SOFF;
    const bool wave_node = UROOT && kBlock && uroot && lst >= 6;
    const unsigned node0u = wave_node ? __builtin_amdgcn_readfirstlane(node0) : node0;
#pragma unroll
    for (int q = 0; q < EPT; q++) {
      if (q & half) continue;
      unsigned node;
      if constexpr (kBlock) {
        node = node0u + (unsigned)(q >> (LOG_EPT - h));
      } else {
        const int pos = map(base + (q << lst));
        node = (1u << s) + ((unsigned)pos >> (LOGN - s));
      }
      uint64_t wu, wup, wv, wvp;
      if (wave_node)
        ld_root2_u(tg, node, wu, wup, wv, wvp);
      else
        ld_root2(tg, node, wu, wup, wv, wvp);
      cp_gi &A = x[q];
      cp_gi &B = x[q + half];
      const uint64_t dr = A.re + p - B.re, di = A.imag + p - B.imag;
      A.re = cp_red2(A.re + B.re, p);
      A.imag = cp_red2(A.imag + B.imag, p);
      if (wave_node)
        cp_shoup2(dr, wu, wup, di, wv, wvp, p, B.re, B.imag);
      else
        cp_shoup2v(dr, wu, wup, di, wv, wvp, p, B.re, B.imag);
    }
    if (CP_STAGE_BARRIER) __builtin_amdgcn_sched_barrier(0);
  }
}

// Split-basis inverse passes over local stages [LSB, LSE) in reverse;
// registers start in the last-pass layout and end in the first-pass layout
// (the exchanges of dit_passes).  AFTER_DIF: the last LDS access of this
// workgroup was the last exchange of dif_passes over the same stages, which
// loaded the registers' layout with the map this sequence's first exchange
// stores with, so the barrier before those stores can go as well.
// `hook` runs after the last LDS read (before the last pass); `early` runs
// before the full pass of loop index EARLY_AT (counting down to 0, the last
// full pass): register loads it starts are covered by that pass and the
// ones after it, which must make no vector-memory loads of their own (wave-
// uniform roots: gsi_pass_rt at lst >= 6).
template <int LOGN, int LOGM, int LSB, int LSE, int SOFF, bool AFTER_DIF = false,
          int EARLY_AT = -1, class Map, class TW, class Hook = NoHook, class Early = NoHook>
__device__ __forceinline__ void gsi_passes(cp_gi (&x)[EPT], cp_gi *lds, int t,
                                           TW __restrict__ tg, uint64_t p, const Map &map,
                                           const Hook &hook = Hook(),
                                           const Early &early = Early()) {
  using PP = PassPlan<LSB, LSE>;
  static_assert(EARLY_AT < 0 || EARLY_AT < PP::NF - 1, "the early hook runs before a loop pass");
  constexpr LdsMap MP = lds_map_for(LOGM - PP::LSF - LOG_EPT);
  t = cp_opaque(t);
  if constexpr (PP::NF > 0) gsi_pass<LOGN, LOGM, PP::LSL, LOG_EPT, SOFF>(x, t, tg, p, map);
  bool prev_local = false, prev_pad = false;
  // map of the previous exchange of this sequence (none yet)
  LdsMap prev_m = AFTER_DIF && PP::NF >= 2
                      ? lds_map_for(LOGM - (PP::LSF + LOG_EPT * (PP::NF - 2)) - 2 * LOG_EPT)
                      : LdsMap{-1, -1};
#pragma unroll  // as in dif_passes
  for (int i = PP::NF - 2; i >= 0; i--) {
    const int ls0 = PP::LSF + LOG_EPT * i;
    const LdsMap m = lds_map_for(LOGM - ls0 - 2 * LOG_EPT);
    const bool local = wave_local(LOGM - ls0 - LOG_EPT), pad = m.A != 0;
    if ((local && prev_local && pad == prev_pad) || same_map(m, prev_m)) wave_sync(); else __syncthreads();
    lds_store_rt<LOGM>(lds, x, t, ls0 + LOG_EPT, m);
    if (local) wave_sync(); else __syncthreads();
    lds_load_rt<LOGM>(lds, x, t, ls0, m);
    if (i == EARLY_AT) early();
    if (PP::R0 == 0 && i == 0) hook();
    gsi_pass_rt<LOGN, LOGM, SOFF, (EARLY_AT >= 0)>(x, t, tg, p, map, ls0, i <= EARLY_AT);
    prev_local = local;
    prev_pad = pad;
    prev_m = m;
  }
  if constexpr (PP::R0 > 0) {
    if constexpr (PP::NF > 0) {
      if (same_map(MP, prev_m)) wave_sync(); else __syncthreads();
      lds_store_rt<LOGM>(lds, x, t, PP::LSF, MP);
      __syncthreads();
      lds_load<LOGM, LSB, PP::R0, MP.A, MP.S>(lds, x, t);
    }
    hook();
    gsi_pass<LOGN, LOGM, LSB, PP::R0, SOFF>(x, t, tg, p, map);
  }
}

// Minimum waves per SIMD asked of a one-transform-per-workgroup kernel: at
// N = 4096 slots (512 threads, 72 KiB of LDS) two workgroups share a CU only
// if each wave keeps to 128 VGPRs.
template <int LOGN>
constexpr int fused_min_waves() { return LOGN == 12 ? 4 : 1; }

// Residue polynomials per workgroup for a single-kernel transform of 2^LOGN
// slots: at least 256 threads per workgroup.
template <int LOGN>
constexpr int rows_for() {
  return ((1 << LOGN) / EPT) >= 256 ? 1 : 256 / ((1 << LOGN) / EPT);
}

// ---- Operand copies into LDS by the DMA path --------------------------------
// global_load_lds_dwordx4 writes a wave's 64 lanes x 16 bytes to LDS at
// M0 + 16 * lane and holds no VGPRs while in flight, so a copy can be started
// before a register pass and overlap it.  Each wave copies its threads' EPT
// first-pass slots of one residue polynomial into its own area of EPT * 64
// slots, lane-major (slot i of lane l at i * 64 + l), and only that wave reads
// the area back, after lds_copy_wait (the wave's own memory counter), with no
// workgroup barrier.  `area` must be wave-uniform.
template <int LOGN>
__device__ __forceinline__ void lds_copy_issue(cp_gi *area, const CP_GLOBAL cp_gi *src, int t) {
  typedef __attribute__((address_space(3))) void lds_void;
  t = cp_opaque(t);
#pragma unroll
  for (int i = 0; i < EPT; i++)
    __builtin_amdgcn_global_load_lds((const CP_GLOBAL void *)(src + first_index<LOGN, 0, LOGN>(t, i)),
                                     (lds_void *)(area + i * 64), 16, 0, CP_CPOL_NT);
}
// The compiler does not order LDS reads after these copies by itself: the
// wait is explicit, and the memory clobber keeps every later LDS read below it.
__device__ __forceinline__ void lds_copy_wait() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }
// Makes the values of x count as computed at this point, so that the work
// producing them (a register pass that loads in flight are meant to overlap)
// is scheduled before the volatile statements that follow.
__device__ __forceinline__ void pin_regs(cp_gi (&x)[EPT]) {
  static_assert(EPT == 8, "one operand pair per slot");
  asm volatile(""
               : "+v"(x[0].re), "+v"(x[0].imag), "+v"(x[1].re), "+v"(x[1].imag),
                 "+v"(x[2].re), "+v"(x[2].imag), "+v"(x[3].re), "+v"(x[3].imag),
                 "+v"(x[4].re), "+v"(x[4].imag), "+v"(x[5].re), "+v"(x[5].imag),
                 "+v"(x[6].re), "+v"(x[6].imag), "+v"(x[7].re), "+v"(x[7].imag));
}
__device__ __forceinline__ void lds_copy_read(cp_gi (&x)[EPT], const cp_gi *area, int lane) {
#pragma unroll
  for (int i = 0; i < EPT; i++) x[i] = area[i * 64 + lane];
}

// ---- n = 2^15: one residue product of N = 16384 slots in one workgroup -----
// Forward stage 0 pairs slot g with g + 8192 (twiddle omega^g).  Thread t of a
// 1024-thread workgroup loads slot g = first_index<13,0,13>(t, i) of both
// halves, so stage 0 (and, at the end, its inverse) runs in registers;
// stages 1..13 run on each 8192-slot half in turn through one 144 KiB LDS
// array (the local transform at stage offset 1).  While one half is in the
// passes the other waits in `park`, a per-workgroup global area of 3 x 8192
// slots addressed thread-major (slot i of thread t at i * 1024 + t: every
// access is coalesced and only the owning thread reads it back), so only
// 8 slots per thread are live during the passes, as in the 8192-slot
// kernels.  Park areas: P0 = a's half-0 spectrum, later the half-0 inverse;
// P1 = a's half 1; P2 = b's half 1.
__device__ __forceinline__ void park_st(CP_GLOBAL cp_gi *park, int area, int t,
                                        const cp_gi (&x)[EPT]) {
  t = cp_opaque(t);
#pragma unroll
  for (int i = 0; i < EPT; i++) st_gi(park + (area * EPT + i) * 1024 + t, x[i]);
}
__device__ __forceinline__ void park_ld(const CP_GLOBAL cp_gi *park, int area, int t,
                                        cp_gi (&x)[EPT]) {
  t = cp_opaque(t);
#pragma unroll
  for (int i = 0; i < EPT; i++) x[i] = ld_gi(park + (area * EPT + i) * 1024 + t);
}

// src's half-0 slots of this thread (slot first_index<13,0,13>(t, i)), for
// stage0_16k with PRE
__device__ __forceinline__ void ld_half0_16k(cp_gi (&x)[EPT], const CP_GLOBAL cp_gi *src, int t) {
  t = cp_opaque(t);
#pragma unroll
  for (int i = 0; i < EPT; i++) x[i] = ld_gi_nt(src + first_index<13, 0, 13>(t, i));
}

// src's half-1 slots of this thread (slot first_index<13,0,13>(t, i) + 8192)
__device__ __forceinline__ void ld_half1_16k(cp_gi (&x)[EPT], const CP_GLOBAL cp_gi *src, int t) {
  t = cp_opaque(t);
#pragma unroll
  for (int i = 0; i < EPT; i++) x[i] = ld_gi_nt(src + first_index<13, 0, 13>(t, i) + 8192);
}

// conversion and twist (split basis) + forward stage 0 of src; half 0 -> x0,
// half 1 -> x1.  PRE = 1: x0 already holds src's half-0 slots (ld_half0_16k);
// PRE = 2: x0 and x1 hold both halves (nothing is loaded).
template <int PRE = 0>
__device__ __forceinline__ void stage0_16k(cp_gi (&x0)[EPT], cp_gi (&x1)[EPT],
                                           const CP_GLOBAL cp_gi *src, int t,
                                           const CP_GLOBAL uint64_t *twf,
                                           const CP_GLOBAL uint64_t *stw, uint64_t J, uint64_t Jp,
                                           uint64_t p) {
  constexpr int H = 8192;
  t = cp_opaque(t);
#pragma unroll
  for (int i = 0; i < EPT; i++) {
    const int g = first_index<13, 0, 13>(t, i);
    cp_gi x = PRE ? x0[i] : ld_gi_nt(src + g), y = PRE == 2 ? x1[i] : ld_gi_nt(src + g + H);
    split_twist2(x, y, stw, g, g + H, J, Jp, p);
    uint64_t w, wp;
    ld_tw(twf, (unsigned)g, w, wp);
    x0[i].re = cp_red2(x.re + y.re, p);
    x0[i].imag = cp_red2(x.imag + y.imag, p);
    cp_shoup2v(x.re + p - y.re, w, wp, x.imag + p - y.imag, w, wp, p, x1[i].re, x1[i].imag);
    __builtin_amdgcn_sched_barrier(0);
  }
}

// split-tree inverse stage 0 (root r_0^-1 of node 1, the same for every
// slot) of slots g and g + 8192 and the conversion back to coefficients
// (the product's second operand carried h: split_back_h2), written to out;
// sk = the residue's split constants.  NEXT: once slot i is written, y[i]
// takes slot i of nxt's half 1 (ld_half1_16k's layout).
template <bool NEXT = false>
__device__ __forceinline__ void istage0_16k(CP_GLOBAL cp_gi *out, const cp_gi (&x)[EPT],
                                            cp_gi (&y)[EPT], int t,
                                            const CP_GLOBAL uint64_t *root,
                                            const uint64_t *sk, uint64_t p,
                                            const CP_GLOBAL cp_gi *nxt = nullptr) {
  constexpr int H = 8192;
  t = cp_opaque(t);
  uint64_t wu, wup, wv, wvp;
  ld_root2_u(root, 1u, wu, wup, wv, wvp);
  const uint64_t J = sk[0], Jp = sk[1];
#pragma unroll
  for (int i = 0; i < EPT; i++) {
    const int g = first_index<13, 0, 13>(t, i);
    cp_gi s, dd;
    s.re = cp_red2(x[i].re + y[i].re, p);
    s.imag = cp_red2(x[i].imag + y[i].imag, p);
    cp_shoup2(x[i].re + p - y[i].re, wu, wup, x[i].imag + p - y[i].imag, wv, wvp, p, dd.re,
              dd.imag);
    split_back_h2(s, dd, J, Jp, p);
    st_gi_nt(out + g, s);
    st_gi_nt(out + g + H, dd);
    if (NEXT) y[i] = ld_gi_nt(nxt + g + H);
    __builtin_amdgcn_sched_barrier(0);
  }
}

// `park` must hold 3 * 8192 slots for this workgroup.  Operand and park
// reads overlap passes that make no vector-memory loads of their own:
// - a forward sequence's last pass (uniform twiddles): a park read into the
//   register set that is free, or b's half 0;
// - an inverse sequence's last full pass and its last one-stage pass
//   (wave-uniform roots): a park read.
// (b's half 1 is read at b's stage 0: an LDS copy of it during a's last
// forward pass, next to the register reads of b's half 0, is waited for as
// soon as those registers are first moved, early in that pass.)
// STREAM (one workgroup walking several units): on entry a's half 0 is being
// copied into the wave areas and y holds a's half 1; before returning, the
// unit starts the same for nxt, the next unit's a (the workgroup's last unit
// passes its own a: the copy and loads are then unused, but unconditional,
// so that the wait counts the compiler derives for the park reads do not
// have to cover a path without them).
template <bool STREAM = false>
__device__ __forceinline__ void polymul_16k_unit(CP_GLOBAL cp_gi *c, const CP_GLOBAL cp_gi *a,
                                                 const CP_GLOBAL cp_gi *b, CP_GLOBAL cp_gi *park,
                                                 cp_gi *lds, const cp_prime &pr,
                                                 const CP_GLOBAL uint64_t *twf,
                                                 const CP_GLOBAL uint64_t *root,
                                                 const CP_GLOBAL uint64_t *stw,
                                                 const CP_GLOBAL uint64_t *stwh,
                                                 const uint64_t *sk, cp_gi (&x)[EPT],
                                                 cp_gi (&y)[EPT],
                                                 const CP_GLOBAL cp_gi *nxt = nullptr) {
  constexpr int H = 8192;
  const int t = (int)threadIdx.x, lane = t & 63;
  cp_gi *area = lds + __builtin_amdgcn_readfirstlane(t >> 6) * (EPT * 64);
  const uint64_t p = pr.p;
  const uint64_t J = sk[0], Jp = sk[1];
  // a: spectrum halves to P0, P1
  if constexpr (STREAM) {
    lds_copy_wait();
    lds_copy_read(x, area, lane);
    stage0_16k<2>(x, y, a, t, twf, stw, J, Jp, p);
  } else {
    stage0_16k(x, y, a, t, twf, stw, J, Jp, p);
  }
  park_st(park, 1, t, y);
  dif_passes<14, 13, 0, 13, 1>(x, lds, t, twf, p, MapBlock{0},
                               [&] { park_ld(park, 1, t, y); });  // a's half 1 -> y
  park_st(park, 0, t, x);
  dif_passes<14, 13, 0, 13, 1>(y, lds, t, twf, p, MapBlock{H},
                               [&] { ld_half0_16k(x, b, t); });  // b's half 0 -> x
  park_st(park, 1, t, y);
  // b: per half, spectrum, product with a's (P0 / P1), inverse stages 13..1;
  // half 0's result waits in P0
  stage0_16k<1>(x, y, b, t, twf, stwh, J, Jp, p);
  park_st(park, 2, t, y);
  dif_passes<14, 13, 0, 13, 1>(x, lds, t, twf, p, MapBlock{0},
                               [&] { park_ld(park, 0, t, y); });  // a's half 0 -> y
#pragma unroll
  for (int i = 0; i < EPT; i++) x[i] = split_mul(x[i], y[i], pr);
  gsi_passes<14, 13, 0, 13, 1, true, 0>(x, lds, t, root, p, MapBlock{0}, NoHook(),
                                        [&] { park_ld(park, 2, t, y); });  // b's half 1 -> y
  park_st(park, 0, t, x);
  dif_passes<14, 13, 0, 13, 1>(y, lds, t, twf, p, MapBlock{H},
                               [&] { park_ld(park, 1, t, x); });  // a's half 1 -> x
#pragma unroll
  for (int i = 0; i < EPT; i++) y[i] = split_mul(y[i], x[i], pr);
  gsi_passes<14, 13, 0, 13, 1, true, 0>(
      y, lds, t, root, p, MapBlock{H},
      [&] {
        if constexpr (STREAM) {  // the next a's half 0 -> the wave areas
          // x's park reads (a full pass ago) are waited for here, before the
          // copy: a wait for a load issued ahead of an LDS copy is compiled
          // as a wait for the copy too
          pin_regs(x);
          __syncthreads();
          lds_copy_issue<13>(area, nxt, t);
          __builtin_amdgcn_sched_barrier(0);
        }
      },
      [&] { park_ld(park, 0, t, x); });  // half 0's result -> x
  istage0_16k<STREAM>(c, x, y, t, root, sk, p, nxt);  // STREAM: the next a's half 1 -> y
}
__device__ __forceinline__ void polymul_16k_unit(CP_GLOBAL cp_gi *c, const CP_GLOBAL cp_gi *a,
                                                 const CP_GLOBAL cp_gi *b, CP_GLOBAL cp_gi *park,
                                                 cp_gi *lds, const cp_prime &pr,
                                                 const CP_GLOBAL uint64_t *twf,
                                                 const CP_GLOBAL uint64_t *root,
                                                 const CP_GLOBAL uint64_t *stw,
                                                 const CP_GLOBAL uint64_t *stwh,
                                                 const uint64_t *sk) {
  cp_gi x[EPT], y[EPT];
  polymul_16k_unit<false>(c, a, b, park, lds, pr, twf, root, stw, stwh, sk, x, y);
}
