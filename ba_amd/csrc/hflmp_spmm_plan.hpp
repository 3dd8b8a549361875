// Row-block layout for Y = A X over a row-major block of k <= 64 vectors,
// and the device walk that applies it.
//
// Rows are grouped into blocks, each the rows of up to 8 runs of
// consecutive rows taken from consecutive "lines" of L rows: block b of
// line group g (lines g nl .. g nl + nl - 1) and segment c holds, from each
// of its nl lines, the rows at positions c a .. c a + a - 1.  With L the
// line length of a structured mesh's numbering this is an a x nl patch of
// the mesh; with nl = 1 and L = nrow it is a run of a consecutive rows.
// For each block the distinct columns of its entries are listed once, in
// ascending order (ucol[uoff[b] .. uoff[b+1])), and every entry j of the
// CSR arrays gets the position of its column in that list (lidx[j]).  A
// workgroup loads a block's X rows into LDS once -- each a contiguous run
// of k values -- and every entry of the block then reads its X row from
// LDS.  A 4 x 4 patch of the 2D Q2 pattern has 16 rows over at most 64
// distinct columns (4 staged X rows per row), where a run of 8 consecutive
// rows has about 55 (7 per row); neighbouring blocks are walked in order
// by one XCD, so the X rows they share stay in its L2.
//
// Each column c of a row is summed in ascending storage order from zero
// (s += a_ij * x_j, no contraction), the order of the single-vector CSR
// loop (CPUsimple_csr_lmatrix.cc:76-111), so column c of the product is
// bit-identical to A times column c.  The layout shares row_ptr and the
// values with the CSR matrix; lidx replaces the column array.
#pragma once

#include "hflmp_internal.hpp"

// Limits of the pipelined walk: distinct columns and entries of one block.
// A wave holds the next block's X rows in registers: 20 of them in FP64,
// 32 in FP32.
__host__ __device__ constexpr int plan_max_cols(int vb) { return vb == 8 ? 80 : 128; }
constexpr int kPlanMaxEntries = 1024;  // 4 per thread

// Block geometry: lines of L rows, blocks of nl runs (one per line) of at
// most a rows; nseg blocks per line group.
struct PlanGeom {
  int nrow, L, a, nl, nseg;
};

constexpr int kPlanMaxRuns = 8;

inline PlanGeom plan_geom(int nrow, int L, int a, int nl) {
  return PlanGeom{nrow, L, a, nl, (L + a - 1) / a};
}
inline int plan_nblk(const PlanGeom &g) {
  const long long lines = ((long long)g.nrow + g.L - 1) / g.L;
  return (int)((lines + g.nl - 1) / g.nl * g.nseg);
}

// Run t (< nl) of block b: rows [r0, r0 + n), clipped to the matrix.
struct PlanRun {
  int r0, n;
};
__host__ __device__ __forceinline__ PlanRun plan_run(const PlanGeom &g, int b, int t) {
  const int grp = b / g.nseg, c = b - grp * g.nseg;
  const long long s = ((long long)grp * g.nl + t) * g.L + (long long)c * g.a;
  const int w = g.a < g.L - c * g.a ? g.a : g.L - c * g.a;
  const long long lo = s < g.nrow ? s : g.nrow;
  const long long hi = s + w < g.nrow ? s + w : g.nrow;
  return PlanRun{(int)lo, (int)(hi - lo)};
}

struct hflmp_spmm_plan {
  int device;
  int nrow, ncol, nnz;
  int L, a, nl;     // block geometry (PlanGeom)
  int R;            // rows per block, at most (a nl)
  int nblk;
  int umax;         // most distinct columns of one block
  int emax;         // most entries of one block
  int nrhs, value_bytes;  // the products the block size was chosen for
  long long nuniq;  // total of the per-block column lists
  int *uoff;        // nblk + 1 offsets into ucol
  int *ucol;        // each block's distinct columns, ascending
  int *lidx;        // per entry: position of its column in its block's list
};

struct PlanView {
  PlanGeom g;
  int nnz, R, nblk, umax, emax;
  const int *__restrict__ uoff;
  const int *__restrict__ ucol;
  const int *__restrict__ lidx;
};

inline PlanView plan_view(const hflmp_spmm_plan *p) {
  return PlanView{plan_geom(p->nrow, p->L, p->a, p->nl), p->nnz, p->R, p->nblk, p->umax,
                  p->emax, p->uoff, p->ucol, p->lidx};
}

// LDS of the walk: the X rows of one block, then two buffers of the
// block's description: entries (values, column positions), column list,
// local row pointers and global row numbers.
__host__ __device__ inline size_t plan_meta_bytes(int R, int umax, int emax, int vb) {
  const size_t b =
      (size_t)emax * (4 + vb) + (size_t)umax * 4 + (size_t)(R + 1) * 4 + (size_t)R * 4;
  return (b + 15) & ~size_t(15);
}
__host__ __device__ inline size_t plan_lds_bytes(int R, int umax, int emax, int k, int vb) {
  const size_t x = ((size_t)umax * k * vb + 15) & ~size_t(15);
  return x + 2 * plan_meta_bytes(R, umax, emax, vb);
}

// Workgroups of a walk: as many as are resident at once on the CUs (the
// kernel's occupancy with `lds` dynamic bytes on top of its static LDS and
// registers; at most 8 per CU), a multiple of the XCD count, and no more
// than the blocks rounded up to that multiple.  A grid beyond what is
// resident would run a second partial round of workgroups.
inline int plan_grid(hflmp_ctx *ctx, const void *kern, int nblk, size_t lds) {
  const int cus = ctx->num_cus > 0 ? ctx->num_cus : 256;
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kBlock, lds) != hipSuccess)
    per_cu = lds > 0 ? (int)(160 * 1024 / lds) : 8;
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
  int *rp;  // local row pointers of the block's rows (R + 1), from 0
  int *gr;  // global row of each of the block's rows (R)
};

template <typename T>
__device__ __forceinline__ PlanMeta<T> plan_meta_at(unsigned char *base, const PlanView &pv) {
  PlanMeta<T> m;
  m.v = reinterpret_cast<T *>(base);
  m.li = reinterpret_cast<int *>(base + (size_t)pv.emax * sizeof(T));
  m.uc = m.li + pv.emax;
  m.rp = m.uc + pv.umax;
  m.gr = m.rp + pv.R + 1;
  return m;
}

// A block's header, one register: lane t < 8 holds rp at the start of run
// t, lane 8 + t rp at its end (runs t >= nl: rp[0] twice), lanes 16 and 17
// uoff[b] and uoff[b + 1].  Loaded a stage before the block's description,
// as a vector load, so that nothing waits on it in the stage that issues it.
__device__ __forceinline__ int plan_hdr_load(int b, const PlanView &pv,
                                             const int *__restrict__ rp) {
  const int lane = threadIdx.x & (kWave - 1);
  const int t = lane & 7;
  const int *p = rp;
  if (lane < 16 && t < pv.g.nl) {
    const PlanRun run = plan_run(pv.g, b, t);
    p = rp + run.r0 + (lane < 8 ? 0 : run.n);
  } else if (lane == 16 || lane == 17) {
    p = pv.uoff + b + (lane - 16);
  }
  return *p;
}

// The uniform values of a block: its number, entries, column list and rows.
struct PlanHdr {
  int b, hreg, ne, u0, U, nr;
};

__device__ __forceinline__ PlanHdr plan_hdr(int hreg, int b, const PlanView &pv) {
  PlanHdr h;
  h.b = b;
  h.hreg = hreg;
  int ne = 0, nr = 0;
  for (int t = 0; t < pv.g.nl; ++t) {
    ne += __builtin_amdgcn_readlane(hreg, 8 + t) - __builtin_amdgcn_readlane(hreg, t);
    nr += plan_run(pv.g, b, t).n;
  }
  h.ne = ne;
  h.nr = nr;
  h.u0 = __builtin_amdgcn_readlane(hreg, 16);
  h.U = __builtin_amdgcn_readlane(hreg, 17) - h.u0;
  return h;
}

// A block's description on its way from memory to LDS: thread t holds
// entries t, t+256, ... (at most 4) of the block's runs laid end to end,
// column t of the list, and the local row pointer and global row of row t.
// Every load is issued (addresses past the block's end are clamped into
// it), so the number of loads in flight does not depend on the block.
template <typename T>
struct PlanMetaRegs {
  int li[4];
  T v[4];
  int uc, rp, gr;
};

template <typename T>
__device__ __forceinline__ void plan_meta_load(PlanMetaRegs<T> &m, const PlanHdr &h,
                                               const PlanView &pv, const int *__restrict__ rp,
                                               const T *__restrict__ val) {
  const int t = threadIdx.x;
  // Local entry e lies in the last run that starts at or before it, and
  // has CSR index e + base for that run's base; local row t likewise has
  // global row t + rbase and its first entry at local index rp[row] + ebase.
  // (Runs are in row order, and an empty run starts where the next one does.)
  int base[4] = {0, 0, 0, 0};
  int rbase = 0, ebase = 0;
  int eacc = 0, racc = 0;
  for (int r = 0; r < pv.g.nl; ++r) {
    const int j = __builtin_amdgcn_readlane(h.hreg, r);
    const int c = __builtin_amdgcn_readlane(h.hreg, 8 + r) - j;
    const PlanRun run = plan_run(pv.g, h.b, r);
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (t + q * kBlock >= eacc) base[q] = j - eacc;
    if (t >= racc) {
      rbase = run.r0 - racc;
      ebase = eacc - j;
    }
    eacc += c;
    racc += run.n;
  }
  // rows past the block's last take its last row (loaded, never stored)
  const int tr = t < h.nr ? t : (h.nr > 0 ? h.nr - 1 : 0);
  m.gr = h.nr > 0 ? tr + rbase : 0;
  m.rp = h.nr > 0 ? rp[m.gr] + ebase : 0;
  m.uc = pv.ucol[h.u0 + (t < h.U ? t : (h.U > 0 ? h.U - 1 : 0))];
  // entries past the block's last read its first run's first entry
  const int j0 = __builtin_amdgcn_readlane(h.hreg, 0);
  const int jpad = j0 < pv.nnz ? j0 : pv.nnz - 1;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int e = t + q * kBlock;
    const int j = e < h.ne ? e + base[q] : jpad;
    m.li[q] = __builtin_nontemporal_load(pv.lidx + j);
    m.v[q] = __builtin_nontemporal_load(val + j);
  }
}

template <typename T>
__device__ __forceinline__ void plan_meta_store(const PlanMetaRegs<T> &m, const PlanHdr &h,
                                                const PlanMeta<T> &d) {
  const int t = threadIdx.x;
  if (t < h.nr) {
    d.rp[t] = m.rp;
    d.gr[t] = m.gr;
  }
  if (t == h.nr) d.rp[t] = h.ne;
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
// pre(i) is called for row i before its sum (a value the epilogue needs,
// loaded early so that its latency hides behind the row's entries); epi(i,
// c, s, u) gets row i's sum s for column c = lane (c >= k: a copy of
// column k - 1, not to be stored) and u = pre(i).
// KC > 0: k is KC, known at compile time (the X-row offsets in LDS are
// then shifts); KC = 0: k from the argument.
template <typename T, int KC, typename Pre, typename Epi>
__device__ __forceinline__ void plan_walk(const PlanView &pv, int k_arg,
                                          const int *__restrict__ rp, const T *__restrict__ val,
                                          const T *__restrict__ X, unsigned char *lds,
                                          Pre &&pre, Epi &&epi) {
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
      const int a = md.rp[r], e = md.rp[r + 1];
      const int gi = md.gr[r];
      const T u = pre(gi);
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
      epi(gi, lane, s, u);
    }
    __syncthreads();
    H0 = H1;
    H1 = H2;
  }
}
