// Discrete Galois Transform kernels for gfx950.
//
// What is computed (the reference's DGTEngine::execute_dgt,
// src/cuda/dgt.cu:681-1094): for each residue, A_n = a_n + i a_{n+N} is
// twisted by xi^n (xi^N = i) and transformed with omega = PROOT^((p-1)/N);
// X_k = sum_n A_n xi^n omega^(nk) lands at position bitrev_N(k).  The
// reference's hierarchical Na x Nb form places X_k at
// bitrev_Na(k/Nb) + Na*bitrev_Nb(k%Nb), which is the same position, so one
// radix-2 decimation-in-frequency network in natural order reproduces it.
// The inverse is the decimation-in-time network with omega^-1 followed by the
// untwist xi^-n / N.
//
// How (designed for CDNA): one workgroup owns a whole residue polynomial (up
// to N = 8192 Gaussian slots = 128 KiB of LDS) or a 2^S-column / block slice
// of a larger one.  Each thread keeps 8 slots in VGPRs and runs three radix-2
// stages on them per pass ("radix-8 pass"); passes exchange data through LDS
// with one barrier pair.  Global loads/stores are issued in the natural
// layout (lane-contiguous 16-byte slots).  Twiddle and twist factors are
// resident tables with Shoup quotients, so every constant product costs one
// mulhi and two mullo on the integer pipes.  The last forward stage (and the
// first inverse stage) has twiddle 1 and is done without a multiply.  The
// fused product kernel keeps a's spectrum in registers while b is
// transformed, multiplies slot-wise and runs the inverse without leaving the
// workgroup: HBM traffic is the 24*k*n bytes of a, b and c.
//
// Sizes: N = 16..8192 in one kernel; N = 16384, 32768 in two kernels
// (stages [0,S) on strided columns, then [S, log N) on contiguous blocks).
#include <utility>
#include <vector>

#include "dgt_dev.hpp"

// -------------------------------------------------------------------------
// Single-kernel transform: N = 2^LOGN <= 8192.  ROWS residue polynomials per
// workgroup; unit u = blockIdx.x*ROWS + row is (poly u / nres, residue u % nres).
// -------------------------------------------------------------------------
template <int LOGN, int ROWS, bool FWD>
__global__ void __launch_bounds__(((1 << LOGN) / EPT) * ROWS)
    k_dgt_single(cp_gi *__restrict__ data, int nunits, int nres, int roff,
                 const cp_prime *__restrict__ primes, const uint64_t *__restrict__ tw_all,
                 const uint64_t *__restrict__ twist_all) {
  constexpr int N = 1 << LOGN, T = N / EPT;
  extern __shared__ cp_gi lds_all[];
  // with one row per workgroup, keep the row (hence the residue and every
  // table pointer) wave-uniform so it lives in SGPRs
  const int row = ROWS == 1 ? 0 : (int)(threadIdx.x / T);
  const int t = ROWS == 1 ? (int)threadIdx.x : (int)(threadIdx.x % T);
  cp_gi *lds = lds_all + row * lds_row<LOGN>();
  int unit = blockIdx.x * ROWS + row;
  const bool active = unit < nunits;
  if (!active) unit = nunits - 1;
  int rid = roff + unit % nres;
  if (ROWS == 1) rid = __builtin_amdgcn_readfirstlane(rid);  // uniform: keep in SGPRs
  const uint64_t p = primes[rid].p;
  const uint64_t *tw = tw_all + (size_t)rid * N;          // N/2 pairs
  const uint64_t *tt = twist_all + (size_t)rid * N * 6;
  cp_gi *poly = data + (size_t)unit * N;
  cp_gi x[EPT];
  MapIdentity id;
  if constexpr (FWD) {
#pragma unroll
    for (int i = 0; i < EPT; i++) {
      const int g = first_index<LOGN, 0, LOGN>(t, i);
      x[i] = poly[g];
      twist(x[i], tt, g, p);
    }
    dif_passes<LOGN, LOGN, 0, LOGN, 0>(x, lds, t, tw, p, id);
    using LP = LastPass<LOGN, 0, LOGN>;
    __syncthreads();
    lds_store<LOGN, LP::LS0, LP::RL, LP::MS.A, LP::MS.S>(lds, x, t);
    __syncthreads();
    if (active) {
#pragma unroll
      for (int i = 0; i < EPT; i++) poly[nat_index<LOGN>(t, i)] = lds[lds_phys(nat_index<LOGN>(t, i), LP::MS)];
    }
  } else {
    using LP = LastPass<LOGN, 0, LOGN>;
#pragma unroll
    for (int i = 0; i < EPT; i++) lds[lds_phys(nat_index<LOGN>(t, i), LP::MS)] = poly[nat_index<LOGN>(t, i)];
    __syncthreads();
    lds_load<LOGN, LP::LS0, LP::RL, LP::MS.A, LP::MS.S>(lds, x, t);
    dit_passes<LOGN, LOGN, 0, LOGN, 0>(x, lds, t, tw, p, id);
    if (active) {
#pragma unroll
      for (int i = 0; i < EPT; i++) {
        const int g = first_index<LOGN, 0, LOGN>(t, i);
        twist(x[i], tt, g, p);
        poly[g] = x[i];
      }
    }
  }
}

// -------------------------------------------------------------------------
// Two-kernel transform: S = LOGN - 7 column stages handled as 2^LOGC
// consecutive columns per workgroup, then the remaining 7 stages on
// contiguous blocks of 128 slots, 2^(LOGM-7) blocks per workgroup; both
// kernels work on a local array of 2^LOGM slots (2^(LOGM-3) threads).
// LOGM = 13 (1024 threads) for N = 16384, 32768; LOGM = 9 (64 threads)
// spreads one residue polynomial of N = 1024 .. 32768 slots over
// 2^(LOGN-9) workgroups per kernel, for calls with too few residue
// polynomials to fill the GPU one workgroup each.
// -------------------------------------------------------------------------
template <int LOGN, int LOGM_ = 13>
struct Split {
  static constexpr int LOGB = 7;                 // block size 128
  static constexpr int S = LOGN - LOGB;          // column stages; column length 2^S
  static constexpr int LOGNS = LOGN - S;         // column stride = 2^LOGB
  static constexpr int LOGM = LOGM_;             // local array 2^LOGM slots
  static constexpr int LOGC = LOGM - S;          // columns per workgroup
  static constexpr int WG_PER_POLY = (1 << LOGNS) >> LOGC;
  static constexpr int BLOCKS_PER_WG = (1 << LOGM) >> LOGB;
  static constexpr int THREADS = (1 << LOGM) / EPT;
  static_assert(LOGC >= 0 && LOGC <= LOGNS, "column split does not fit the local array");
};
// log2 of the local array of the wide split (64-thread workgroups)
constexpr int kWideLogM = 9;

// Column stages [0, S): forward (twist first) or inverse (untwist last).
template <int LOGN, bool FWD, int LOGM_ = 13>
__global__ void __launch_bounds__((1 << LOGM_) / EPT)
    k_dgt_cols(cp_gi *__restrict__ data, int nunits, int nres, int roff,
               const cp_prime *__restrict__ primes, const uint64_t *__restrict__ tw_all,
               const uint64_t *__restrict__ twist_all) {
  using SP = Split<LOGN, LOGM_>;
  constexpr int N = 1 << LOGN, LOGM = SP::LOGM;
  extern __shared__ cp_gi lds[];
  const int t = threadIdx.x;
  const int unit = blockIdx.x / SP::WG_PER_POLY;
  const int col0 = (blockIdx.x % SP::WG_PER_POLY) << SP::LOGC;
  const int rid = roff + unit % nres;
  const uint64_t p = primes[rid].p;
  const uint64_t *tw = tw_all + (size_t)rid * N;
  const uint64_t *tt = twist_all + (size_t)rid * N * 6;
  cp_gi *poly = data + (size_t)unit * N;
  MapColumns<SP::LOGC, SP::LOGNS> map{col0};
  cp_gi x[EPT];
  // local stages [0, S) of the 2^13 local array are global stages [0, S)
  if constexpr (FWD) {
#pragma unroll
    for (int i = 0; i < EPT; i++) {
      const int g = map(first_index<LOGM, 0, SP::S>(t, i));
      x[i] = poly[g];
      twist(x[i], tt, g, p);
    }
    dif_passes<LOGN, LOGM, 0, SP::S, 0>(x, lds, t, tw, p, map);
    using LP = LastPass<LOGM, 0, SP::S>;
    __syncthreads();
    lds_store<LOGM, LP::LS0, LP::RL, LP::MS.A, LP::MS.S>(lds, x, t);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < EPT; i++) poly[map(nat_index<LOGM>(t, i))] = lds[lds_phys(nat_index<LOGM>(t, i), LP::MS)];
  } else {
    using LP = LastPass<LOGM, 0, SP::S>;
#pragma unroll
    for (int i = 0; i < EPT; i++) lds[lds_phys(nat_index<LOGM>(t, i), LP::MS)] = poly[map(nat_index<LOGM>(t, i))];
    __syncthreads();
    lds_load<LOGM, LP::LS0, LP::RL, LP::MS.A, LP::MS.S>(lds, x, t);
    dit_passes<LOGN, LOGM, 0, SP::S, 0>(x, lds, t, tw, p, map);
#pragma unroll
    for (int i = 0; i < EPT; i++) {
      const int g = map(first_index<LOGM, 0, SP::S>(t, i));
      twist(x[i], tt, g, p);
      poly[g] = x[i];
    }
  }
  (void)nunits;
}

// One 16384-slot residue polynomial (n = 2^15) per 1024-thread workgroup, in
// place, in the standard layout: the values of the split launch (k_dgt_cols +
// k_dgt_blocks; the same network, the arithmetic exact) with one read and one
// write of the polynomial and one thread-major park round trip of a half.
// Stage 0 pairs slot g with g + 8192 in registers (g = first_index<13,0,13>
// (t, i): the thread's slots are t modulo 1024), and the two 8192-slot halves
// take stages 1..13 in turn through the LDS, each stored (forward) or loaded
// (inverse) in natural order through the LDS.  The parked half waits in the
// polynomial itself:
// - forward: half 1 at H + i * 1024 + t, slots this thread has read;
// - inverse: half 0 at i * 1024 + t, after every thread has loaded half 0
//   into the LDS (the barriers of its passes come between).
template <bool FWD>
__global__ void __launch_bounds__(1024)
    k_dgt_single16(cp_gi *__restrict__ data, int nres, int roff,
                   const cp_prime *__restrict__ primes, const uint64_t *__restrict__ tw_all,
                   const uint64_t *__restrict__ twist_all) {
  constexpr int N = 16384, H = 8192, LOGM = 13;
  using LP = LastPass<LOGM, 0, LOGM>;
  extern __shared__ cp_gi lds[];
  const int t = (int)threadIdx.x;
  const int unit = (int)blockIdx.x;
  const int rid = __builtin_amdgcn_readfirstlane(roff + unit % nres);
  const uint64_t p = primes[rid].p;
  const uint64_t *tw = tw_all + (size_t)rid * N;
  const uint64_t *tt = twist_all + (size_t)rid * N * 6;
  CP_GLOBAL cp_gi *gp = (CP_GLOBAL cp_gi *)(data + (size_t)unit * N);
  cp_gi x[EPT], y[EPT];
  if constexpr (FWD) {
    {
      const int ta = cp_opaque(t);
#pragma unroll
      for (int i = 0; i < EPT; i++) {  // twist, then global stage 0 (twiddle omega^g)
        const int g = first_index<LOGM, 0, LOGM>(ta, i);
        cp_gi a = ld_gi(gp + g), b = ld_gi(gp + g + H);
        twist(a, tt, g, p);
        twist(b, tt, g + H, p);
        uint64_t w, wp;
        ld_tw(tw, (unsigned)g, w, wp);
        x[i].re = cp_red2(a.re + b.re, p);
        x[i].imag = cp_red2(a.imag + b.imag, p);
        cp_shoup2v(a.re + p - b.re, w, wp, a.imag + p - b.imag, w, wp, p, y[i].re, y[i].imag);
      }
    }
    park_st(gp + H, 0, t, y);
#pragma unroll
    for (int h = 0; h < 2; h++) {
      if (h == 1) park_ld(gp + H, 0, t, x);
      dif_passes<14, LOGM, 0, LOGM, 1>(x, lds, t, tw, p, MapBlock{h * H});
      __syncthreads();
      lds_store<LOGM, LP::LS0, LP::RL, LP::MS.A, LP::MS.S>(lds, x, t);
      __syncthreads();
#pragma unroll
      for (int i = 0; i < EPT; i++)
        st_gi(gp + h * H + nat_index<LOGM>(t, i), lds[lds_phys(nat_index<LOGM>(t, i), LP::MS)]);
    }
  } else {
    // both halves pass through x (half 0 then parked, read back into y)
#pragma unroll
    for (int h = 0; h < 2; h++) {
      if (h == 1) __syncthreads();  // the LDS reads of half 0's passes are done
#pragma unroll
      for (int i = 0; i < EPT; i++)
        lds[lds_phys(nat_index<LOGM>(t, i), LP::MS)] = ld_gi(gp + h * H + nat_index<LOGM>(t, i));
      __syncthreads();
      lds_load<LOGM, LP::LS0, LP::RL, LP::MS.A, LP::MS.S>(lds, x, t);
      dit_passes<14, LOGM, 0, LOGM, 1>(x, lds, t, tw, p, MapBlock{h * H});
      if (h == 0) park_st(gp, 0, t, x);
    }
    // no global store comes between the park's store and this read, so the
    // compiler could otherwise hoist the read above half 1's passes
    asm volatile("" ::: "memory");
    park_ld(gp, 0, t, y);
    const int ta = cp_opaque(t);
#pragma unroll
    for (int i = 0; i < EPT; i++) {  // global stage 0 (twiddle omega^-g), then the untwist
      const int g = first_index<LOGM, 0, LOGM>(ta, i);
      uint64_t w, wp, tr, ti;
      ld_tw(tw, (unsigned)g, w, wp);
      cp_shoup2v(x[i].re, w, wp, x[i].imag, w, wp, p, tr, ti);  // half 1 (slot g + H)
      cp_gi a, b;
      b.re = cp_red2(y[i].re + p - tr, p);
      b.imag = cp_red2(y[i].imag + p - ti, p);
      a.re = cp_red2(y[i].re + tr, p);
      a.imag = cp_red2(y[i].imag + ti, p);
      twist(a, tt, g, p);
      twist(b, tt, g + H, p);
      st_gi(gp + g, a);
      st_gi(gp + g + H, b);
      __builtin_amdgcn_sched_barrier(0);  // one slot's table reads in flight at a time
    }
  }
}

// Block stages [S, LOGN): each workgroup row is one 128-slot block.
template <int LOGN, bool FWD, int LOGM_ = 13>
__global__ void __launch_bounds__((1 << LOGM_) / EPT)
    k_dgt_blocks(cp_gi *__restrict__ data, int nunits, int nres, int roff,
                 const cp_prime *__restrict__ primes, const uint64_t *__restrict__ tw_all) {
  using SP = Split<LOGN, LOGM_>;
  constexpr int N = 1 << LOGN, LOGB = SP::LOGB, B = 1 << LOGB, TB = B / EPT;
  extern __shared__ cp_gi lds_all[];
  const int row = threadIdx.x / TB, t = threadIdx.x % TB;
  const int gblk = blockIdx.x * SP::BLOCKS_PER_WG + row;  // global block over all units
  const int unit = gblk >> (LOGN - LOGB);
  const int blk = gblk & ((1 << (LOGN - LOGB)) - 1);
  const int rid = roff + unit % nres;
  const uint64_t p = primes[rid].p;
  const uint64_t *tw = tw_all + (size_t)rid * N;
  cp_gi *lds = lds_all + row * lds_row<LOGB>();
  cp_gi *poly = data + (size_t)unit * N;
  MapBlock map{blk << LOGB};
  cp_gi x[EPT];
  using LP = LastPass<LOGB, 0, LOGB>;
  if constexpr (FWD) {
#pragma unroll
    for (int i = 0; i < EPT; i++) x[i] = poly[map(first_index<LOGB, 0, LOGB>(t, i))];
    dif_passes<LOGN, LOGB, 0, LOGB, SP::S>(x, lds, t, tw, p, map);
    __syncthreads();
    lds_store<LOGB, LP::LS0, LP::RL, LP::MS.A, LP::MS.S>(lds, x, t);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < EPT; i++) poly[map(nat_index<LOGB>(t, i))] = lds[lds_phys(nat_index<LOGB>(t, i), LP::MS)];
  } else {
#pragma unroll
    for (int i = 0; i < EPT; i++) lds[lds_phys(nat_index<LOGB>(t, i), LP::MS)] = poly[map(nat_index<LOGB>(t, i))];
    __syncthreads();
    lds_load<LOGB, LP::LS0, LP::RL, LP::MS.A, LP::MS.S>(lds, x, t);
    dit_passes<LOGN, LOGB, 0, LOGB, SP::S>(x, lds, t, tw, p, map);
#pragma unroll
    for (int i = 0; i < EPT; i++) poly[map(first_index<LOGB, 0, LOGB>(t, i))] = x[i];
  }
  (void)nunits;
}

// -------------------------------------------------------------------------
// Fused coefficient-domain product c = a*b, N <= 8192, in the split basis
// (dgt_dev.hpp): conversion and twist of a, forward network, the same for b,
// two products mod p per slot, the split-tree inverse and the conversion back,
// in one workgroup.  The DGT-domain values of a and b never leave it, so the
// basis does not show; the coefficients are the exact negacyclic product.
// For N = 4096 (512 threads, 72 KiB of LDS) two workgroups fit on a CU only
// if each wave keeps to 128 VGPRs, so that size asks for 4 waves per SIMD
// (fused_min_waves).
// -------------------------------------------------------------------------

template <int LOGN, int ROWS>
__global__ void __launch_bounds__(((1 << LOGN) / EPT) * ROWS, fused_min_waves<LOGN>())
    k_polymul_fused(cp_gi *__restrict__ out, const cp_gi *__restrict__ a,
                    const cp_gi *__restrict__ b, int nunits, int nres,
                    const cp_prime *__restrict__ primes, const uint64_t *__restrict__ twf_all,
                    const uint64_t *__restrict__ root_all, const uint64_t *__restrict__ stw_all,
                    const uint64_t *__restrict__ stwh_all, const uint64_t *__restrict__ sk_all) {
  constexpr int N = 1 << LOGN, T = N / EPT;
  extern __shared__ cp_gi lds_all[];
  // with one row per workgroup, keep the row (hence the residue and every
  // table pointer) wave-uniform so it lives in SGPRs
  const int row = ROWS == 1 ? 0 : (int)(threadIdx.x / T);
  const int t = ROWS == 1 ? (int)threadIdx.x : (int)(threadIdx.x % T);
  cp_gi *lds = lds_all + row * lds_row<LOGN>();
  int unit = blockIdx.x * ROWS + row;
  const bool active = unit < nunits;
  if (!active) unit = nunits - 1;
  int rid = unit % nres;
  if (ROWS == 1) rid = __builtin_amdgcn_readfirstlane(rid);  // uniform: keep in SGPRs
  const cp_prime pr = primes[rid];
  const uint64_t p = pr.p;
  const uint64_t *twf = twf_all + (size_t)rid * N;
  const uint64_t *rootp = root_all + (size_t)rid * N * 4, *stwp = stw_all + (size_t)rid * N * 4,
                 *stwhp = stwh_all + (size_t)rid * N * 4;
  const uint64_t *sk = sk_all + (size_t)rid * 8;
  if (ROWS == 1) {
    rootp = uniform_ptr(rootp);
    stwp = uniform_ptr(stwp);
    stwhp = uniform_ptr(stwhp);
    a = uniform_ptr(a + (size_t)unit * N);
    b = uniform_ptr(b + (size_t)unit * N);
    out = uniform_ptr(out + (size_t)unit * N);
  } else {
    a += (size_t)unit * N;
    b += (size_t)unit * N;
    out += (size_t)unit * N;
  }
  const CP_GLOBAL uint64_t *root = (const CP_GLOBAL uint64_t *)rootp;
  const CP_GLOBAL uint64_t *stw = (const CP_GLOBAL uint64_t *)stwp;
  const CP_GLOBAL uint64_t *stwh = (const CP_GLOBAL uint64_t *)stwhp;
  const uint64_t J = sk[0], Jp = sk[1];
  MapIdentity id;
  cp_gi xa[EPT], xb[EPT];
  // a: load, convert and twist, transform; b is loaded only afterwards so
  // that a single operand is live during each forward transform.  The
  // scheduling barriers keep the twist-table loads of one slot pair from being
  // hoisted above the previous pair's products (which would spill).
#pragma unroll
  for (int i = 0; i < EPT; i++)
    xa[i] = ld_gi_nt((const CP_GLOBAL cp_gi *)a + first_index<LOGN, 0, LOGN>(t, i));
#pragma unroll
  for (int i = 0; i < EPT; i += 2) {
    split_twist2(xa[i], xa[i + 1], stw, first_index<LOGN, 0, LOGN>(t, i),
                 first_index<LOGN, 0, LOGN>(t, i + 1), J, Jp, p);
    __builtin_amdgcn_sched_barrier(0);
  }
  // b is loaded during a's last forward pass (its twiddles are scalar loads)
  dif_passes<LOGN, LOGN, 0, LOGN, 0>(xa, lds, t, twf, p, id, [&] {
    const int tq = cp_opaque(t);
#pragma unroll
    for (int i = 0; i < EPT; i++)
      xb[i] = ld_gi_nt((const CP_GLOBAL cp_gi *)b + first_index<LOGN, 0, LOGN>(tq, i));
  });
  // an opaque copy of the thread index: the address arithmetic of the later
  // loads and stores cannot be hoisted above this point (and held in VGPRs
  // across the transform).
  int tb = t;
  asm volatile("" : "+v"(tb));
#pragma unroll
  for (int i = 0; i < EPT; i += 2) {
    split_twist2(xb[i], xb[i + 1], stwh, first_index<LOGN, 0, LOGN>(tb, i),
                 first_index<LOGN, 0, LOGN>(tb, i + 1), J, Jp, p);
    __builtin_amdgcn_sched_barrier(0);
  }
  dif_passes<LOGN, LOGN, 0, LOGN, 0>(xb, lds, t, twf, p, id);
#pragma unroll
  for (int i = 0; i < EPT; i++) xa[i] = split_mul_paired(xa[i], xb[i], pr);
  gsi_passes<LOGN, LOGN, 0, LOGN, 0, true>(xa, lds, t, root, p, id);
  int to = t;
  asm volatile("" : "+v"(to));
  if (active) {
#pragma unroll
    for (int i = 0; i < EPT; i += 2) {
      split_back_h2(xa[i], xa[i + 1], J, Jp, p);
      st_gi_nt((CP_GLOBAL cp_gi *)out + first_index<LOGN, 0, LOGN>(to, i), xa[i]);
      st_gi_nt((CP_GLOBAL cp_gi *)out + first_index<LOGN, 0, LOGN>(to, i + 1), xa[i + 1]);
      __builtin_amdgcn_sched_barrier(0);
    }
  }
}

// Fused block stages for N > 8192: forward blocks of a and b, product, inverse blocks.
template <int LOGN>
__global__ void __launch_bounds__(1024)
    k_polymul_blocks(cp_gi *__restrict__ a, const cp_gi *__restrict__ b, int nres,
                     const cp_prime *__restrict__ primes, const uint64_t *__restrict__ twf_all,
                     const uint64_t *__restrict__ twi_all) {
  using SP = Split<LOGN>;
  constexpr int N = 1 << LOGN, LOGB = SP::LOGB, B = 1 << LOGB, TB = B / EPT;
  extern __shared__ cp_gi lds_all[];
  const int row = threadIdx.x / TB, t = threadIdx.x % TB;
  const int gblk = blockIdx.x * SP::BLOCKS_PER_WG + row;
  const int unit = gblk >> (LOGN - LOGB);
  const int blk = gblk & ((1 << (LOGN - LOGB)) - 1);
  const int rid = unit % nres;
  const cp_prime pr = primes[rid];
  const uint64_t p = pr.p;
  const uint64_t *twf = twf_all + (size_t)rid * N, *twi = twi_all + (size_t)rid * N;
  cp_gi *lds = lds_all + row * lds_row<LOGB>();
  cp_gi *pa = a + (size_t)unit * N;
  const cp_gi *pb = b + (size_t)unit * N;
  MapBlock map{blk << LOGB};
  cp_gi xa[EPT], xb[EPT];
#pragma unroll
  for (int i = 0; i < EPT; i++) {
    xa[i] = pa[map(first_index<LOGB, 0, LOGB>(t, i))];
    xb[i] = pb[map(first_index<LOGB, 0, LOGB>(t, i))];
  }
  dif_passes<LOGN, LOGB, 0, LOGB, SP::S>(xa, lds, t, twf, p, map);
  dif_passes<LOGN, LOGB, 0, LOGB, SP::S>(xb, lds, t, twf, p, map);
#pragma unroll
  for (int i = 0; i < EPT; i++) xa[i] = cp_gimul(xa[i], xb[i], pr);
  dit_passes<LOGN, LOGB, 0, LOGB, SP::S>(xa, lds, t, twi, p, map);
#pragma unroll
  for (int i = 0; i < EPT; i++) pa[map(first_index<LOGB, 0, LOGB>(t, i))] = xa[i];
}

// Product for N = 16384 slots (n = 2^15): one 1024-thread workgroup per
// residue product (polymul_16k_unit), persistent over the batch; each
// workgroup owns 3 * 8192 park slots.  Streaming: each unit starts loading
// the next unit's a (half 0 into the LDS wave areas, half 1 into registers)
// during its last inverse passes and its stores, so the workgroups' operand
// reads overlap their arithmetic instead of arriving in one burst.
__global__ void __launch_bounds__(1024)
    k_polymul_16k(cp_gi *__restrict__ out, const cp_gi *__restrict__ a,
                  const cp_gi *__restrict__ b, int nunits, int nres,
                  const cp_prime *__restrict__ primes, const uint64_t *__restrict__ twf_all,
                  const uint64_t *__restrict__ root_all, const uint64_t *__restrict__ stw_all,
                  const uint64_t *__restrict__ stwh_all, const uint64_t *__restrict__ sk_all,
                  cp_gi *__restrict__ park_all) {
  constexpr int N = 16384;
  extern __shared__ cp_gi lds[];
  CP_GLOBAL cp_gi *park = (CP_GLOBAL cp_gi *)(park_all + (size_t)blockIdx.x * 3 * (N / 2));
  const int t = (int)threadIdx.x;
  int unit = blockIdx.x;
  if (unit >= nunits) return;  // uniform
  cp_gi x[EPT], y[EPT];
  {  // the first unit's a: half 0 into the wave areas, half 1 into y
    const CP_GLOBAL cp_gi *a0 = (const CP_GLOBAL cp_gi *)uniform_ptr(a + (size_t)unit * N);
    lds_copy_issue<13>(lds + __builtin_amdgcn_readfirstlane(t >> 6) * (EPT * 64), a0, t);
    ld_half1_16k(y, a0, t);
  }
  for (; unit < nunits; unit += gridDim.x) {  // uniform trip count per workgroup
    const int rid = unit % nres;
    const int next = unit + (int)gridDim.x;
    const cp_prime pr = primes[rid];
    const size_t off = (size_t)unit * N;
    polymul_16k_unit<true>(
        (CP_GLOBAL cp_gi *)(out + off), (const CP_GLOBAL cp_gi *)(a + off),
        (const CP_GLOBAL cp_gi *)uniform_ptr(b + off), park, lds, pr,
        (const CP_GLOBAL uint64_t *)(twf_all + (size_t)rid * N),
        (const CP_GLOBAL uint64_t *)(root_all + (size_t)rid * N * 4),
        (const CP_GLOBAL uint64_t *)(stw_all + (size_t)rid * N * 4),
        (const CP_GLOBAL uint64_t *)(stwh_all + (size_t)rid * N * 4), sk_all + (size_t)rid * 8, x,
        y, (const CP_GLOBAL cp_gi *)uniform_ptr(a + (size_t)(next < nunits ? next : unit) * N));
  }
}

// Persistent form of the batched product with one residue polynomial per
// workgroup (n/2 = 2^10 .. 2^13 slots).  The grid holds as many workgroups as
// the GPU runs at once; each walks the units blockIdx.x, + gridDim.x, ...
// Operand loads start while the previous transform runs its last register
// pass: b's straight into registers during a's last forward pass; the next
// unit's a as an LDS copy (lds_copy_issue: no registers are free there)
// during the last inverse pass, the conversion back and the stores of this
// unit.  One-shot workgroups all load their operands at the same moments and
// leave the CUs waiting on bursts of HBM reads.  Results are the same as
// k_polymul_fused<LOGN, 1> (same passes, same arithmetic).
template <int LOGN>
__global__ void __launch_bounds__((1 << LOGN) / EPT)
    k_polymul_stream(cp_gi *__restrict__ out, const cp_gi *__restrict__ a,
                     const cp_gi *__restrict__ b, int nunits, int nres,
                     const cp_prime *__restrict__ primes, const uint64_t *__restrict__ twf_all,
                     const uint64_t *__restrict__ root_all, const uint64_t *__restrict__ stw_all,
                     const uint64_t *__restrict__ stwh_all, const uint64_t *__restrict__ sk_all) {
  constexpr int N = 1 << LOGN, T = N / EPT;
  static_assert(T % 64 == 0 && PassPlan<0, LOGN>::NF > 0, "whole waves, a register last pass");
  static_assert(T * EPT <= lds_row<LOGN>(), "the copy areas fit the exchange array");
  extern __shared__ cp_gi lds[];
  const int t = (int)threadIdx.x, lane = t & 63;
  cp_gi *area = lds + __builtin_amdgcn_readfirstlane(t >> 6) * (EPT * 64);
  MapIdentity id;
  int unit = blockIdx.x;
  if (unit >= nunits) return;  // uniform
  lds_copy_issue<LOGN>(area, (const CP_GLOBAL cp_gi *)uniform_ptr(a + (size_t)unit * N), t);
  for (; unit < nunits; unit += gridDim.x) {  // uniform trip count per workgroup
    const int rid = __builtin_amdgcn_readfirstlane(unit % nres);
    const cp_prime pr = primes[rid];
    const uint64_t p = pr.p;
    const UTab twt = utab(uniform_ptr(twf_all + (size_t)rid * N), N * 16);
    const UTab *twf = &twt;
    const CP_GLOBAL uint64_t *root =
        (const CP_GLOBAL uint64_t *)uniform_ptr(root_all + (size_t)rid * N * 4);
    const CP_GLOBAL uint64_t *stw =
        (const CP_GLOBAL uint64_t *)uniform_ptr(stw_all + (size_t)rid * N * 4);
    const CP_GLOBAL uint64_t *stwh =
        (const CP_GLOBAL uint64_t *)uniform_ptr(stwh_all + (size_t)rid * N * 4);
    const uint64_t *sk = sk_all + (size_t)rid * 8;
    const uint64_t J = sk[0], Jp = sk[1];
    const int next = unit + (int)gridDim.x;
    cp_gi xa[EPT], xb[EPT];
    // a: from its LDS copy; convert and twist; transform.  b is loaded into
    // registers at the start of a's last forward pass (which has no other
    // vector-memory loads: its twiddles are uniform, scalar loads).
    lds_copy_wait();
    lds_copy_read(xa, area, lane);
    {
      const int ta = cp_opaque(t);
#pragma unroll
      for (int i = 0; i < EPT; i += 2) {
        split_twist2(xa[i], xa[i + 1], stw, first_index<LOGN, 0, LOGN>(ta, i),
                     first_index<LOGN, 0, LOGN>(ta, i + 1), J, Jp, p);
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    const CP_GLOBAL cp_gi *ub = (const CP_GLOBAL cp_gi *)uniform_ptr(b + (size_t)unit * N);
    dif_passes<LOGN, LOGN, 0, LOGN, 0>(xa, lds, t, twf, p, id, [&] {
      const int tq = cp_opaque(t);
#pragma unroll
      for (int i = 0; i < EPT; i++) xb[i] = ld_gi_nt(ub + first_index<LOGN, 0, LOGN>(tq, i));
    });
    pin_regs(xa);
    {
      const int tb = cp_opaque(t);
#pragma unroll
      for (int i = 0; i < EPT; i += 2) {
        split_twist2(xb[i], xb[i + 1], stwh, first_index<LOGN, 0, LOGN>(tb, i),
                     first_index<LOGN, 0, LOGN>(tb, i + 1), J, Jp, p);
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    dif_passes<LOGN, LOGN, 0, LOGN, 0>(xb, lds, t, twf, p, id);
#pragma unroll
    for (int i = 0; i < EPT; i++) xa[i] = split_mul_paired(xa[i], xb[i], pr);
    gsi_passes<LOGN, LOGN, 0, LOGN, 0, true>(xa, lds, t, root, p, id, [&] {
      __syncthreads();
      if (next < nunits)
        lds_copy_issue<LOGN>(area, (const CP_GLOBAL cp_gi *)uniform_ptr(a + (size_t)next * N),
                             t);
    });
    CP_GLOBAL cp_gi *uo = (CP_GLOBAL cp_gi *)uniform_ptr(out + (size_t)unit * N);
    const int to = cp_opaque(t);
#pragma unroll
    for (int i = 0; i < EPT; i += 2) {
      split_back_h2(xa[i], xa[i + 1], J, Jp, p);
      st_gi_nt(uo + first_index<LOGN, 0, LOGN>(to, i), xa[i]);
      st_gi_nt(uo + first_index<LOGN, 0, LOGN>(to, i + 1), xa[i + 1]);
      __builtin_amdgcn_sched_barrier(0);
    }
  }
}

// -------------------------------------------------------------------------
// Host side
// -------------------------------------------------------------------------

template <class K>
static int set_lds(K kernel, size_t bytes) {
  if (bytes > 64 * 1024)
    CP_HIP(hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)bytes));
  return CUPOLY_OK;
}

template <int LOGN>
static int setup_size() {
  int rc;
  if constexpr (LOGN <= 13) {
    constexpr int R = rows_for<LOGN>();
    const size_t lds = (size_t)R * lds_row<LOGN>() * sizeof(cp_gi);
    if ((rc = set_lds(k_dgt_single<LOGN, R, true>, lds))) return rc;
    if ((rc = set_lds(k_dgt_single<LOGN, R, false>, lds))) return rc;
    if ((rc = set_lds(k_polymul_fused<LOGN, R>, lds))) return rc;
    if constexpr (LOGN >= 10 && LOGN <= 12) {
      if ((rc = set_lds(k_polymul_fused<LOGN, 1>, lds_row<LOGN>() * sizeof(cp_gi)))) return rc;
      if ((rc = set_lds(k_polymul_fused<LOGN, 2>, 2 * lds_row<LOGN>() * sizeof(cp_gi)))) return rc;
    }
    if constexpr (LOGN >= 10 && LOGN <= 11)
      if ((rc = set_lds(k_polymul_fused<LOGN, 4>, 4 * lds_row<LOGN>() * sizeof(cp_gi)))) return rc;
    if constexpr (LOGN == 10)
      if ((rc = set_lds(k_polymul_fused<LOGN, 8>, 8 * lds_row<LOGN>() * sizeof(cp_gi)))) return rc;
    if constexpr (LOGN >= 10)
      if ((rc = set_lds(k_polymul_stream<LOGN>, lds_row<LOGN>() * sizeof(cp_gi)))) return rc;
  } else {
    const size_t lds = (size_t)lds_row<13>() * sizeof(cp_gi);
    if ((rc = set_lds(k_dgt_cols<LOGN, true>, lds))) return rc;
    if ((rc = set_lds(k_dgt_cols<LOGN, false>, lds))) return rc;
    if ((rc = set_lds(k_dgt_blocks<LOGN, true>, lds))) return rc;
    if ((rc = set_lds(k_dgt_blocks<LOGN, false>, lds))) return rc;
    if ((rc = set_lds(k_polymul_blocks<LOGN>, lds))) return rc;
    if constexpr (LOGN == 14) {
      if ((rc = set_lds(k_dgt_single16<true>, lds))) return rc;
      if ((rc = set_lds(k_dgt_single16<false>, lds))) return rc;
    }
    if constexpr (LOGN == 14)
      if ((rc = set_lds(k_polymul_16k, lds))) return rc;
  }
  return CUPOLY_OK;
}

int cp_dgt_setup(cupoly_engine *e) {
  switch (e->logng) {
    case 4: return setup_size<4>();
    case 5: return setup_size<5>();
    case 6: return setup_size<6>();
    case 7: return setup_size<7>();
    case 8: return setup_size<8>();
    case 9: return setup_size<9>();
    case 10: return setup_size<10>();
    case 11: return setup_size<11>();
    case 12: return setup_size<12>();
    case 13: return setup_size<13>();
    case 14: return setup_size<14>();
    case 15: return setup_size<15>();
    default: return cp_fail(CUPOLY_ERR_UNSUPPORTED, "DGT is not working for such N");
  }
}

template <int LOGN, int LOGM>
static void launch_split(cupoly_ctx *c, cp_gi *data, int roff, int nres, int nunits, int dir) {
  using SP = Split<LOGN, LOGM>;
  cupoly_engine *e = c->e;
  const size_t lds_c = (size_t)lds_row<LOGM>() * sizeof(cp_gi);
  const size_t lds_b = (size_t)SP::BLOCKS_PER_WG * lds_row<SP::LOGB>() * sizeof(cp_gi);
  const int gcols = nunits * SP::WG_PER_POLY;
  const int gblks = (nunits << (LOGN - SP::LOGB)) / SP::BLOCKS_PER_WG;
  const dim3 th(SP::THREADS);
  if (dir == CUPOLY_FORWARD) {
    hipLaunchKernelGGL((k_dgt_cols<LOGN, true, LOGM>), dim3(gcols), th, lds_c, c->stream, data,
                       nunits, nres, roff, e->d_primes, e->d_tw_fwd, e->d_twist_fwd);
    hipLaunchKernelGGL((k_dgt_blocks<LOGN, true, LOGM>), dim3(gblks), th, lds_b, c->stream, data,
                       nunits, nres, roff, e->d_primes, e->d_tw_fwd);
  } else {
    hipLaunchKernelGGL((k_dgt_blocks<LOGN, false, LOGM>), dim3(gblks), th, lds_b, c->stream,
                       data, nunits, nres, roff, e->d_primes, e->d_tw_inv);
    hipLaunchKernelGGL((k_dgt_cols<LOGN, false, LOGM>), dim3(gcols), th, lds_c, c->stream, data,
                       nunits, nres, roff, e->d_primes, e->d_tw_inv, e->d_twist_inv);
  }
}

// One transform launch form per call: N <= 8192 slots in one workgroup per
// residue polynomial (k_dgt_single); N = 16384 in one workgroup per residue
// polynomial as well (k_dgt_single16) once those cover 5/8 of the CUs;
// otherwise N = 16384, 32768 in the 1024-thread split; when that would give
// fewer workgroups than CUs and N >= 1024, the
// 64-thread split, which spreads each residue polynomial over 2^(LOGN-9)
// workgroups per kernel (the reference's two-pass form, dgt.cu:125-160,
// whose passes run on Na * nres blocks).
template <int LOGN>
static int dgt_launch_t(cupoly_ctx *c, cp_gi *data, int roff, int nres, int npoly, int dir) {
  cupoly_engine *e = c->e;
  const int nunits = nres * npoly;
  if (nunits <= 0) return CUPOLY_OK;
  if constexpr (LOGN >= 10) {
    int wgs = nunits;
    if constexpr (LOGN > 13) wgs *= Split<LOGN>::WG_PER_POLY;
    if (wgs < e->ncu) {
      launch_split<LOGN, kWideLogM>(c, data, roff, nres, nunits, dir);
      CP_LAUNCH_CHECK();
      return CUPOLY_OK;
    }
  }
  if constexpr (LOGN <= 13) {
    constexpr int R = rows_for<LOGN>(), T = (1 << LOGN) / EPT;
    const size_t lds = (size_t)R * lds_row<LOGN>() * sizeof(cp_gi);
    const int grid = (nunits + R - 1) / R;
    if (dir == CUPOLY_FORWARD)
      hipLaunchKernelGGL((k_dgt_single<LOGN, R, true>), dim3(grid), dim3(T * R), lds, c->stream,
                         data, nunits, nres, roff, e->d_primes, e->d_tw_fwd, e->d_twist_fwd);
    else
      hipLaunchKernelGGL((k_dgt_single<LOGN, R, false>), dim3(grid), dim3(T * R), lds, c->stream,
                         data, nunits, nres, roff, e->d_primes, e->d_tw_inv, e->d_twist_inv);
  } else if (LOGN == 14 && 8 * nunits >= 5 * e->ncu) {
    // one workgroup per residue polynomial once they cover 5/8 of the CUs
    const size_t lds = (size_t)lds_row<13>() * sizeof(cp_gi);
    if (dir == CUPOLY_FORWARD)
      hipLaunchKernelGGL(k_dgt_single16<true>, dim3(nunits), dim3(1024), lds, c->stream, data, nres,
                         roff, e->d_primes, e->d_tw_fwd, e->d_twist_fwd);
    else
      hipLaunchKernelGGL(k_dgt_single16<false>, dim3(nunits), dim3(1024), lds, c->stream, data,
                         nres, roff, e->d_primes, e->d_tw_inv, e->d_twist_inv);
  } else {
    launch_split<LOGN, 13>(c, data, roff, nres, nunits, dir);
  }
  CP_LAUNCH_CHECK();
  return CUPOLY_OK;
}

int cp_dgt_launch(cupoly_ctx *c, cp_gi *data, int roff, int nres, int npoly, int dir) {
  switch (c->e->logng) {
    case 4: return dgt_launch_t<4>(c, data, roff, nres, npoly, dir);
    case 5: return dgt_launch_t<5>(c, data, roff, nres, npoly, dir);
    case 6: return dgt_launch_t<6>(c, data, roff, nres, npoly, dir);
    case 7: return dgt_launch_t<7>(c, data, roff, nres, npoly, dir);
    case 8: return dgt_launch_t<8>(c, data, roff, nres, npoly, dir);
    case 9: return dgt_launch_t<9>(c, data, roff, nres, npoly, dir);
    case 10: return dgt_launch_t<10>(c, data, roff, nres, npoly, dir);
    case 11: return dgt_launch_t<11>(c, data, roff, nres, npoly, dir);
    case 12: return dgt_launch_t<12>(c, data, roff, nres, npoly, dir);
    case 13: return dgt_launch_t<13>(c, data, roff, nres, npoly, dir);
    case 14: return dgt_launch_t<14>(c, data, roff, nres, npoly, dir);
    case 15: return dgt_launch_t<15>(c, data, roff, nres, npoly, dir);
    default: return cp_fail(CUPOLY_ERR_UNSUPPORTED, "DGT is not working for such N");
  }
}

template <int LOGN, int R>
static void launch_fused(cupoly_ctx *c, cp_gi *out, const cp_gi *a, const cp_gi *b, int nunits) {
  static_assert(R * (1 << LOGN) <= 8192, "a workgroup holds at most 8192 slots");
  constexpr int T = (1 << LOGN) / EPT;
  cupoly_engine *e = c->e;
  const size_t lds = (size_t)R * lds_row<LOGN>() * sizeof(cp_gi);
  hipLaunchKernelGGL((k_polymul_fused<LOGN, R>), dim3((nunits + R - 1) / R), dim3(T * R), lds,
                     c->stream, out, a, b, nunits, e->k, e->d_primes, e->d_tw_fwd,
                     e->d_split_root, e->d_split_tw, e->d_split_tw_h, e->d_split_k);
}

// Persistent (streaming) form: as many workgroups as the GPU holds at once.
template <int LOGN>
static int launch_stream(cupoly_ctx *c, cp_gi *out, const cp_gi *a, const cp_gi *b, int nunits) {
  constexpr int T = (1 << LOGN) / EPT;
  cupoly_engine *e = c->e;
  const size_t lds = (size_t)lds_row<LOGN>() * sizeof(cp_gi);
  static int per_cu = 0;  // the same for every gfx950 device
  if (per_cu == 0) {
    int n = 0;
    CP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_polymul_stream<LOGN>, T, lds));
    per_cu = n > 0 ? n : 1;
  }
  const long slots = (long)per_cu * e->ncu;
  const int grid = nunits < slots ? nunits : (int)slots;
  hipLaunchKernelGGL(k_polymul_stream<LOGN>, dim3(grid), dim3(T), lds, c->stream, out, a, b,
                     nunits, e->k, e->d_primes, e->d_tw_fwd, e->d_split_root, e->d_split_tw,
                     e->d_split_tw_h, e->d_split_k);
  return CUPOLY_OK;
}

// Rows (residue polynomials) per workgroup the batched product can use at
// this size besides the built-in rows_for: n = 2^11 .. 2^13 only.
static bool fused_rows_ok(int logng, int rows) {
  if (logng < 10 || logng > 12) return false;
  return (rows == 1 || rows == 2 || rows == 4 || rows == 8) && (rows << logng) <= 8192;
}

template <int LOGN>
static int polymul_t(cupoly_ctx *c, cp_gi *out, const cp_gi *a, const cp_gi *b, int batch) {
  cupoly_engine *e = c->e;
  const int nres = e->k, nunits = nres * batch;
  if (nunits <= 0) return CUPOLY_OK;
  // n = 2^15: one workgroup per residue product once the batch covers 5/8 of
  // the CUs (measured at k = 4: 40 products 0.26 against 0.29 ms split, 32
  // products 0.25 against 0.17 ms)
  const bool persistent =
      c->tune_persistent ? c->tune_persistent > 0 : 8 * nunits >= 5 * e->ncu;
  if constexpr (LOGN <= 13) {
    // one residue polynomial per workgroup by default from n = 2^11 (at
    // n = 2^11, k = 4, batch 8192: 1.03 ms against 1.08 ms with two)
    const int rows = fused_rows_ok(LOGN, c->tune_rows) ? c->tune_rows
                     : LOGN == 10                      ? 1
                                                       : rows_for<LOGN>();
    // the streaming form: by default at n = 2^14 once the batch has more
    // residue products than CUs (k = 4, batch 1024: 1.342 against 1.384 ms
    // one-shot); at n = 2^11 .. 2^13, where several one-shot workgroups share
    // a CU, it is slower (1.08 / 1.14 / 1.22 against 1.03 / 1.10 / 1.18 ms)
    // and runs only when asked for
    if constexpr (LOGN >= 10) {
      if (rows == 1 && c->tune_persistent >= 0 &&
          (c->tune_persistent > 0 || (LOGN == 13 && nunits > e->ncu))) {
        int rc = launch_stream<LOGN>(c, out, a, b, nunits);
        if (rc) return rc;
        CP_LAUNCH_CHECK();
        return CUPOLY_OK;
      }
    }
    if constexpr (LOGN >= 10 && LOGN <= 12) {
      if (rows == 1) launch_fused<LOGN, 1>(c, out, a, b, nunits);
      else if (rows == 2) launch_fused<LOGN, 2>(c, out, a, b, nunits);
      else if constexpr (LOGN <= 11) {
        if (rows == 4) launch_fused<LOGN, 4>(c, out, a, b, nunits);
        else if constexpr (LOGN == 10) launch_fused<LOGN, 8>(c, out, a, b, nunits);
      }
    } else {
      launch_fused<LOGN, rows_for<LOGN>()>(c, out, a, b, nunits);
    }
    CP_LAUNCH_CHECK();
  } else if (LOGN == 14 && persistent) {
    // persistent: one workgroup per CU, each with its own 384 KiB park area.
    // By default smaller batches keep the split below, which spreads each
    // product over twice as many workgroups.
    const int grid = nunits < e->ncu ? nunits : e->ncu;
    void *ws;
    int rc = cp_workspace(c, (size_t)grid * 3 * (1 << (LOGN - 1)) * sizeof(cp_gi), &ws);
    if (rc) return rc;
    const size_t lds = (size_t)lds_row<13>() * sizeof(cp_gi);
    hipLaunchKernelGGL(k_polymul_16k, dim3(grid), dim3(1024), lds, c->stream, out, a, b, nunits,
                       nres, e->d_primes, e->d_tw_fwd, e->d_split_root, e->d_split_tw,
                       e->d_split_tw_h, e->d_split_k, (cp_gi *)ws);
    CP_LAUNCH_CHECK();
  } else {
    // out <- a, scratch <- b (caller inputs are not modified); transforms in place
    using SP = Split<LOGN>;
    const size_t bytes = (size_t)nunits * (1 << LOGN) * sizeof(cp_gi);
    const size_t lds = (size_t)lds_row<13>() * sizeof(cp_gi);
    // b first: out may be b itself
    cp_gi *tb = nullptr;
    CP_HIP(hipMallocAsync((void **)&tb, bytes, c->stream));
    // from here on every exit releases tb
    hipError_t err = hipMemcpyAsync(tb, b, bytes, hipMemcpyDeviceToDevice, c->stream);
    if (err == hipSuccess && out != a)
      err = hipMemcpyAsync(out, a, bytes, hipMemcpyDeviceToDevice, c->stream);
    if (err == hipSuccess) {
      const int gcols = nunits * SP::WG_PER_POLY;
      const int gblks = (nunits << (LOGN - SP::LOGB)) / SP::BLOCKS_PER_WG;
      hipLaunchKernelGGL((k_dgt_cols<LOGN, true>), dim3(gcols), dim3(1024), lds, c->stream, out,
                         nunits, nres, 0, e->d_primes, e->d_tw_fwd, e->d_twist_fwd);
      hipLaunchKernelGGL((k_dgt_cols<LOGN, true>), dim3(gcols), dim3(1024), lds, c->stream, tb,
                         nunits, nres, 0, e->d_primes, e->d_tw_fwd, e->d_twist_fwd);
      hipLaunchKernelGGL((k_polymul_blocks<LOGN>), dim3(gblks), dim3(1024), lds, c->stream, out,
                         tb, nres, e->d_primes, e->d_tw_fwd, e->d_tw_inv);
      hipLaunchKernelGGL((k_dgt_cols<LOGN, false>), dim3(gcols), dim3(1024), lds, c->stream, out,
                         nunits, nres, 0, e->d_primes, e->d_tw_inv, e->d_twist_inv);
      err = hipGetLastError();
    }
    const hipError_t ferr = hipFreeAsync(tb, c->stream);
    if (err != hipSuccess) return cp_check_hip(err, "polymul (split transforms)");
    CP_HIP(ferr);
  }
  return CUPOLY_OK;
}

int cp_polymul_fused_launch(cupoly_ctx *c, cp_gi *out, const cp_gi *a, const cp_gi *b, int batch) {
  switch (c->e->logng) {
    case 4: return polymul_t<4>(c, out, a, b, batch);
    case 5: return polymul_t<5>(c, out, a, b, batch);
    case 6: return polymul_t<6>(c, out, a, b, batch);
    case 7: return polymul_t<7>(c, out, a, b, batch);
    case 8: return polymul_t<8>(c, out, a, b, batch);
    case 9: return polymul_t<9>(c, out, a, b, batch);
    case 10: return polymul_t<10>(c, out, a, b, batch);
    case 11: return polymul_t<11>(c, out, a, b, batch);
    case 12: return polymul_t<12>(c, out, a, b, batch);
    case 13: return polymul_t<13>(c, out, a, b, batch);
    case 14: return polymul_t<14>(c, out, a, b, batch);
    case 15: return polymul_t<15>(c, out, a, b, batch);
    default: return cp_fail(CUPOLY_ERR_UNSUPPORTED, "DGT is not working for such N");
  }
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" int cupoly_dgt(cupoly_ctx *c, cupoly_gi *data, int base, int direction) {
  if (!c || !data) return cp_fail(CUPOLY_ERR_ARG, "dgt: null argument");
  if (base == CUPOLY_TBASE) return CUPOLY_OK;  // dgt.cu:1084-1085
  int off, cnt;
  if (!cp_base_range(c->e, base, &off, &cnt)) return cp_fail(CUPOLY_ERR_UNSUPPORTED, "Unknown base");
  if (direction != CUPOLY_FORWARD && direction != CUPOLY_INVERSE)
    return cp_fail(CUPOLY_ERR_ARG, "dgt: bad direction");
  return cp_dgt_launch(c, (cp_gi *)data, off, cnt, 1, direction);
}

extern "C" int cupoly_polymul_batch(cupoly_ctx *c, cupoly_gi *out, const cupoly_gi *a,
                                    const cupoly_gi *b, int batch) {
  // an empty batch is a no-op (its operand pointers may be null)
  if (c && batch == 0) return CUPOLY_OK;
  if (!c || !out || !a || !b || batch < 0) return cp_fail(CUPOLY_ERR_ARG, "polymul_batch: bad argument");
  if ((long long)batch * c->e->k >= (1LL << 31))
    return cp_fail(CUPOLY_ERR_ARG, "polymul_batch: batch * k must stay below 2^31");
  return cp_polymul_fused_launch(c, (cp_gi *)out, (const cp_gi *)a, (const cp_gi *)b, batch);
}

extern "C" int cupoly_polymul_set_tuning(cupoly_ctx *c, int rows, int persistent) {
  if (!c) return cp_fail(CUPOLY_ERR_ARG, "polymul_set_tuning: null context");
  if (rows != 0 && !fused_rows_ok(c->e->logng, rows))
    return cp_fail(CUPOLY_ERR_ARG, "polymul_set_tuning: rows " + std::to_string(rows) +
                                       " not available at this ring degree");
  if (persistent < -1 || persistent > 1)
    return cp_fail(CUPOLY_ERR_ARG, "polymul_set_tuning: persistent must be -1, 0 or 1");
  if (persistent != 0 && (c->e->logng < 10 || c->e->logng > 14))
    return cp_fail(CUPOLY_ERR_ARG,
                   "polymul_set_tuning: the persistent choice exists for n = 2^11 .. 2^15 only");
  c->tune_rows = rows;
  c->tune_persistent = persistent;
  return CUPOLY_OK;
}

extern "C" int cupoly_polymul_autotune(cupoly_ctx *c, cupoly_gi *out, const cupoly_gi *a,
                                       const cupoly_gi *b, int batch, int reps, int *rows,
                                       int *persistent) {
  if (!c || !out || !a || !b || batch < 1 || reps < 1)
    return cp_fail(CUPOLY_ERR_ARG, "polymul_autotune: bad argument");
  if (out == a || out == b)  // every candidate runs on the same operands
    return cp_fail(CUPOLY_ERR_ARG, "polymul_autotune: out must not be a or b");
  // candidate (rows, persistent) pairs for this ring degree
  std::vector<std::pair<int, int>> cand;
  const int logng = c->e->logng;
  for (int r : {1, 2, 4, 8})
    if (fused_rows_ok(logng, r)) cand.push_back({r, r == 1 ? -1 : 0});
  if (logng >= 10 && logng <= 12) cand.push_back({1, 1});
  if (logng == 13 || logng == 14) cand = {{0, 1}, {0, -1}};
  if (cand.empty()) cand.push_back({0, 0});
  hipEvent_t e0, e1;
  CP_HIP(hipEventCreate(&e0));
  if (hipError_t err = hipEventCreate(&e1); err != hipSuccess) {
    (void)hipEventDestroy(e0);
    return cp_check_hip(err, "hipEventCreate");
  }
  auto hip = [](hipError_t err, const char *what) {
    return err == hipSuccess ? CUPOLY_OK : cp_check_hip(err, what);
  };
  float best = -1.0f;
  std::pair<int, int> pick = cand[0];
  const std::pair<int, int> before{c->tune_rows, c->tune_persistent};
  int rc = CUPOLY_OK;
  for (const auto &cd : cand) {
    c->tune_rows = cd.first;
    c->tune_persistent = cd.second;
    if ((rc = cupoly_polymul_batch(c, out, a, b, batch))) break;  // warm-up
    if ((rc = hip(hipEventRecord(e0, c->stream), "hipEventRecord"))) break;
    for (int i = 0; i < reps && !rc; i++) rc = cupoly_polymul_batch(c, out, a, b, batch);
    if (rc || (rc = hip(hipEventRecord(e1, c->stream), "hipEventRecord")) ||
        (rc = hip(hipEventSynchronize(e1), "hipEventSynchronize")))
      break;
    float ms = 0.0f;
    if ((rc = hip(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime"))) break;
    if (best < 0.0f || ms < best) {
      best = ms;
      pick = cd;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) {  // a failed search leaves the caller's choice in place
    c->tune_rows = before.first;
    c->tune_persistent = before.second;
    return rc;
  }
  c->tune_rows = pick.first;
  c->tune_persistent = pick.second;
  if (rows) *rows = pick.first;
  if (persistent) *persistent = pick.second;
  return cupoly_polymul_batch(c, out, a, b, batch);
}
