// Row-block layout for Y = A X over a row-major block of k <= 64 vectors,
// and the device walk that applies it.
//
// Rows are cut into blocks of R consecutive rows.  For each block the
// distinct columns of its entries are listed once, in ascending order
// (ucol[uoff[b] .. uoff[b+1])), and every entry j of the CSR arrays gets the
// position of its column in that list (lidx[j]).  A workgroup loads a
// block's X rows into LDS once -- each a contiguous run of k values -- and
// every entry of the block then reads its X row from LDS.  For an FE
// stiffness matrix a block of 8 rows of the 2D Q2 pattern has ~25 entries
// per row over ~55 distinct columns, so each X row crosses L2 -> CU about
// 3 times less often than with one gather per entry, and with the blocks of
// one XCD walked in order the X rows shared by neighbouring blocks stay in
// that XCD's L2 between their loads.
//
// Each column c of a row is summed in ascending storage order from zero
// (s += a_ij * x_j, no contraction), the order of the single-vector CSR
// loop (CPUsimple_csr_lmatrix.cc:76-111), so column c of the product is
// bit-identical to A times column c.  The layout shares row_ptr and the
// values with the CSR matrix; lidx replaces the column array.
#pragma once

#include "hflmp_internal.hpp"

// Limits of the pipelined walk: distinct columns and entries of one block.
// A wave holds the next block's X rows in registers: 16 of them in FP64,
// 32 in FP32.
__host__ __device__ constexpr int plan_max_cols(int vb) { return vb == 8 ? 64 : 128; }
constexpr int kPlanMaxEntries = 1024;  // 4 per thread

struct hflmp_spmm_plan {
  int device;
  int nrow, ncol, nnz;
  int R;            // rows per block
  int nblk;         // ceil(nrow / R)
  int umax;         // most distinct columns of one block
  int emax;         // most entries of one block
  int nrhs, value_bytes;  // the products the block size was chosen for
  long long nuniq;  // total of the per-block column lists
  int *uoff;        // nblk + 1 offsets into ucol
  int *ucol;        // each block's distinct columns, ascending
  int *lidx;        // per entry: position of its column in its block's list
};

struct PlanView {
  int nrow, nnz, R, nblk, umax, emax;
  const int *__restrict__ uoff;
  const int *__restrict__ ucol;
  const int *__restrict__ lidx;
};

inline PlanView plan_view(const hflmp_spmm_plan *p) {
  return PlanView{p->nrow, p->nnz, p->R, p->nblk, p->umax, p->emax, p->uoff, p->ucol, p->lidx};
}

// LDS of the walk: the X rows of one block, then two buffers of the
// block's entries (column positions, values), column list and row pointers.
__host__ __device__ inline size_t plan_meta_bytes(int R, int umax, int emax, int vb) {
  const size_t b = (size_t)emax * (4 + vb) + (size_t)umax * 4 + (size_t)(R + 1) * 4;
  return (b + 15) & ~size_t(15);
}
__host__ __device__ inline size_t plan_lds_bytes(int R, int umax, int emax, int k, int vb) {
  const size_t x = ((size_t)umax * k * vb + 15) & ~size_t(15);
  return x + 2 * plan_meta_bytes(R, umax, emax, vb);
}

// Workgroups of a walk: as many as fit on the CUs with `lds` bytes each (at
// most 8 per CU), a multiple of the XCD count, and no more than the blocks
// rounded up to that multiple.
inline int plan_grid(int num_cus, int nblk, size_t lds) {
  const int cus = num_cus > 0 ? num_cus : 256;
  int per_cu = lds > 0 ? (int)(160 * 1024 / lds) : 8;
  if (per_cu > 8) per_cu = 8;
  if (per_cu < 1) per_cu = 1;
  long long g = (long long)per_cu * cus / kXcds * kXcds;
  const long long gb = ((long long)nblk + kXcds - 1) / kXcds * kXcds;
  if (g > gb) g = gb;
  if (g < kXcds) g = kXcds;
  return (int)g;
}

// One block's description in LDS.
template <typename T>
struct PlanMeta {
  int *li;  // entries' column positions
  T *v;     // entries' values
  int *uc;  // distinct columns
  int *rp;  // row pointers of the block's rows (R + 1)
};

template <typename T>
__device__ __forceinline__ PlanMeta<T> plan_meta_at(unsigned char *base, const PlanView &pv) {
  PlanMeta<T> m;
  m.v = reinterpret_cast<T *>(base);
  m.li = reinterpret_cast<int *>(base + (size_t)pv.emax * sizeof(T));
  m.uc = m.li + pv.emax;
  m.rp = m.uc + pv.umax;
  return m;
}

// A block's header: lanes 0..3 of one register hold rp[r0], rp[r0 + nr],
// uoff[b], uoff[b + 1].  Loaded a stage before the block's description,
// as a vector load, so that nothing waits on it in the stage that issues it.
__device__ __forceinline__ int plan_hdr_load(int b, const PlanView &pv,
                                             const int *__restrict__ rp) {
  const int lane = threadIdx.x & (kWave - 1);
  const int r0 = b * pv.R;
  const int r1 = r0 + pv.R < pv.nrow ? r0 + pv.R : pv.nrow;
  const int *p = lane == 1 ? rp + r1 : lane == 2 ? pv.uoff + b : lane == 3 ? pv.uoff + b + 1
                                                                          : rp + r0;
  return *p;
}

// The uniform values of a block, from its header register.
struct PlanHdr {
  int r0, nr, j0, ne, u0, U;
};

__device__ __forceinline__ PlanHdr plan_hdr(int hreg, int b, const PlanView &pv) {
  PlanHdr h;
  h.r0 = b * pv.R;
  h.nr = h.r0 + pv.R < pv.nrow ? pv.R : pv.nrow - h.r0;
  h.j0 = __builtin_amdgcn_readlane(hreg, 0);
  h.ne = __builtin_amdgcn_readlane(hreg, 1) - h.j0;
  h.u0 = __builtin_amdgcn_readlane(hreg, 2);
  h.U = __builtin_amdgcn_readlane(hreg, 3) - h.u0;
  return h;
}

// A block's description on its way from memory to LDS: thread t holds
// entries t, t+256, ... (at most 4), column t of the list and row pointer
// t.  Every load is issued (addresses past the block's end are clamped
// into it), so the number of loads in flight does not depend on the block.
template <typename T>
struct PlanMetaRegs {
  int li[4];
  T v[4];
  int uc, rp;
};

template <typename T>
__device__ __forceinline__ void plan_meta_load(PlanMetaRegs<T> &m, const PlanHdr &h,
                                               const PlanView &pv, const int *__restrict__ rp,
                                               const T *__restrict__ val) {
  const int t = threadIdx.x;
  m.rp = rp[h.r0 + (t < h.nr ? t : h.nr)];
  m.uc = pv.ucol[h.u0 + (t < h.U ? t : (h.U > 0 ? h.U - 1 : 0))];
  const int last = h.ne > 0 ? h.j0 + h.ne - 1 : (h.j0 < pv.nnz ? h.j0 : pv.nnz - 1);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int e = h.j0 + t + q * kBlock;
    const int j = e < last ? e : last;
    m.li[q] = __builtin_nontemporal_load(pv.lidx + j);
    m.v[q] = __builtin_nontemporal_load(val + j);
  }
}

template <typename T>
__device__ __forceinline__ void plan_meta_store(const PlanMetaRegs<T> &m, const PlanHdr &h,
                                                const PlanMeta<T> &d) {
  const int t = threadIdx.x;
  if (t <= h.nr) d.rp[t] = m.rp;
  if (t < h.U) d.uc[t] = m.uc;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int e = t + q * kBlock;
    if (e < h.ne) {
      d.li[e] = m.li[q];
      d.v[e] = m.v[q];
    }
  }
}

// The walk.  XCD x (workgroups b with b % 8 == x) takes the contiguous band
// of blocks [x band, (x+1) band) and its workgroups stride over it, so the
// workgroups of one XCD are on neighbouring blocks at any time.  Per
// workgroup a four-stage pipeline: while block m is computed from LDS, the
// X rows of block m+1 (wave w: rows w, w+4, ...; lane c: value c), the
// description of block m+2 and the header of block m+3 are in flight into
// registers; the next step stores them to LDS after waiting for loads
// issued a whole step before.  Per row, the entries' positions and values
// are read from LDS (the same address across the wave) and their X values
// from the staged rows, eight in flight, added in storage order.
// epi(i, s) gets row i's sum for column `lane` (lanes >= k hold a copy of
// column k - 1 and must not store).
// KC > 0: k is KC, known at compile time (the X-row offsets in LDS are
// then shifts); KC = 0: k from the argument.
template <typename T, int KC, typename Epi>
__device__ __forceinline__ void plan_walk(const PlanView &pv, int k_arg,
                                          const int *__restrict__ rp, const T *__restrict__ val,
                                          const T *__restrict__ X, unsigned char *lds,
                                          Epi &&epi) {
#pragma clang fp contract(off)
  const int k = KC > 0 ? KC : k_arg;
  constexpr int nw = kBlock / kWave;
  constexpr int PFX = plan_max_cols((int)sizeof(T)) / nw;  // X rows per wave
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
  const int lk = lane < k ? lane : k - 1;
  const int per = gridDim.x / kXcds;
  const int xcd = blockIdx.x % kXcds, bx = blockIdx.x / kXcds;
  const int band = (pv.nblk + kXcds - 1) / kXcds;
  const int b_beg = xcd * band + bx;
  const int b_end = (xcd + 1) * band < pv.nblk ? (xcd + 1) * band : pv.nblk;
  if (b_beg >= b_end) return;
  const int nb = (b_end - b_beg + per - 1) / per;  // blocks of this workgroup
  // block m of this workgroup; past the last one, the last one again (its
  // loads are issued and never used)
  auto blk = [&](int m) { return b_beg + (m < nb ? m : nb - 1) * per; };

  T *xs = reinterpret_cast<T *>(lds);
  unsigned char *mb = lds + (((size_t)pv.umax * k * sizeof(T) + 15) & ~size_t(15));
  const size_t msz = plan_meta_bytes(pv.R, pv.umax, pv.emax, (int)sizeof(T));
  // buffer `i` of the descriptions, recomputed from the LDS base each time
  // (an array of the two pointers would hide that they point into LDS)
  auto meta = [&](int i) { return plan_meta_at<T>(mb + (size_t)i * msz, pv); };

  PlanMetaRegs<T> mr;
  T xr[PFX];
  auto load_x = [&](const PlanMeta<T> &m, int U) {
#pragma unroll
    for (int q = 0; q < PFX; ++q) {
      const int u = wid + q * nw;
      if (u < U) xr[q] = X[(size_t)m.uc[u] * k + lk];
    }
  };
  auto store_x = [&](int U) {
    if (lane < k) {
#pragma unroll
      for (int q = 0; q < PFX; ++q) {
        const int u = wid + q * nw;
        if (u < U) xs[u * k + lane] = xr[q];
      }
    }
  };

  // prologue: block 0's description in buffer 0, block 1's in registers,
  // block 0's X rows and block 2's header in flight
  const int h0 = plan_hdr_load(blk(0), pv, rp);
  const int h1 = plan_hdr_load(blk(1), pv, rp);
  PlanHdr H0 = plan_hdr(h0, blk(0), pv);
  plan_meta_load(mr, H0, pv, rp, val);
  plan_meta_store(mr, H0, meta(0));
  PlanHdr H1 = plan_hdr(h1, blk(1), pv);
  plan_meta_load(mr, H1, pv, rp, val);
  __syncthreads();
  load_x(meta(0), H0.U);
  int hn = plan_hdr_load(blk(2), pv, rp);  // header of block m+2
  for (int m = 0; m < nb; ++m) {
    const int buf = m & 1;
    // loads of the previous step: X rows of m, description of m+1, header of m+2
    store_x(H0.U);
    plan_meta_store(mr, H1, meta(buf ^ 1));
    const PlanHdr H2 = plan_hdr(hn, blk(m + 2), pv);
    __syncthreads();
    // this step's loads
    hn = plan_hdr_load(blk(m + 3), pv, rp);
    plan_meta_load(mr, H2, pv, rp, val);
    load_x(meta(buf ^ 1), H1.U);
    // block m from LDS
    const PlanMeta<T> md = meta(buf);
    for (int r = wid; r < H0.nr; r += nw) {
      const int a = md.rp[r] - H0.j0, e = md.rp[r + 1] - H0.j0;
      T s = 0;
      int j = a;
      for (; j + 8 <= e; j += 8) {
        int li[8];
        T vv[8], xv[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          li[q] = md.li[j + q];
          vv[q] = md.v[j + q];
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) xv[q] = xs[li[q] * k + lk];
#pragma unroll
        for (int q = 0; q < 8; ++q) s += vv[q] * xv[q];
      }
      for (; j < e; ++j) s += md.v[j] * xs[md.li[j] * k + lk];
      epi(H0.r0 + r, s);
    }
    __syncthreads();
    H0 = H1;
    H1 = H2;
  }
}
