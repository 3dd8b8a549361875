// Device-resident solve server: many independent FP64 Jacobi-PCG solves of
// different sizes advanced together.  No reference counterpart (the
// reference runs one CG::Solve at a time, linear_solver/cg.h:165-246); each
// system follows that arithmetic with its own scalars and IterateControl
// test (common/iterate_control.cc:80-107).
//
// Memory.  One device arena of the HBM budget is allocated at creation.  A
// request is admitted (first come, first served) when a slot is free and
// its footprint fits in the arena; its CSR arrays, SELL-64 layout, inverse
// diagonal, right-hand side and CG vectors then occupy one arena block
// until its result has been copied back.
//
// Batching.  The rows of every active system are cut into units of 256
// consecutive rows (one workgroup, four SELL-64 slices).  One launch of each
// iteration kernel covers the units of all active systems (unit -> system
// through a device table), and the two scalar steps run one workgroup per
// system.  So an iteration of S systems is five launches, as for one.  A
// system's partial sums are indexed by its own units and added in a fixed
// order, so its iterates do not depend on which other systems share the
// launches: a result is bit-identical alone and in any mix.
//
// Steps.  hflmp_server_step admits what fits, forms the initial residuals
// of new systems, runs `chunk` iterations (captured once per active set as
// a hipGraph and replayed while the set is unchanged), reads the states,
// and retires the systems that stopped (x copied to the caller's buffer).
#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <vector>

#include "hflmp_sell.hpp"
#include "hflmp_solver_common.hpp"

namespace {

using namespace hflmp_solver;

constexpr int kFresh = -2;  // admitted; initial residual not yet formed

struct SysDesc {
  int nrow, nslice, nunits, unit0;
  const long long *soff;
  const int *slen;
  const unsigned char *rlen;
  const int *col;
  const double *sval;
  const double *invd;
  const double *b;
  double *x, *r, *p, *z, *Ap, *part;
  CgState<double> *st;
};

__device__ __forceinline__ SellOp<double, int> sys_op(const SysDesc &d) {
  SellOp<double, int> op;
  op.nrow = d.nrow;
  op.nslice = d.nslice;
  op.soff = d.soff;
  op.slen = d.slen;
  op.rlen = d.rlen;
  op.col = d.col;
  op.val = d.sval;
  return op;
}

// sum of src[stride*u + off], u < count, over one workgroup of kBlock
// threads, eight loads in flight per thread; valid in thread 0
__device__ __forceinline__ double units_sum(const double *__restrict__ src, int count, int stride,
                                            int off) {
  __shared__ double lds[kBlock / kWave];
  double a[8];
#pragma unroll
  for (int f = 0; f < 8; ++f) a[f] = 0;
  for (int u0 = threadIdx.x; u0 < count; u0 += 8 * kBlock) {
    double v[8];
#pragma unroll
    for (int f = 0; f < 8; ++f) {
      const int u = u0 + f * kBlock;
      v[f] = u < count ? src[(size_t)u * stride + off] : 0.0;
    }
#pragma unroll
    for (int f = 0; f < 8; ++f) a[f] += v[f];
  }
  const double t = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
  return block_sum(t, lds);
}

__device__ __forceinline__ void unit_partial2(double a, double b, double *part, int j) {
  __shared__ double lds[2][kBlock / kWave];
  a = wave_sum(a);
  b = wave_sum(b);
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  if (lane == 0) {
    lds[0][wid] = a;
    lds[1][wid] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double sa = 0, sb = 0;
#pragma unroll
    for (int w = 0; w < kBlock / kWave; ++w) {
      sa += lds[0][w];
      sb += lds[1][w];
    }
    part[2 * j] = sa;
    part[2 * j + 1] = sb;
  }
}

// Unit of the launch -> (system descriptor, local unit).  Units of the
// launch are spread over the XCDs in contiguous bands.
__device__ __forceinline__ bool unit_of(const SysDesc *__restrict__ sys,
                                        const int *__restrict__ unit_sys, int U, SysDesc &d,
                                        int &j) {
  const long long u = xcd_tile(blockIdx.x, gridDim.x);
  if (u >= U) return false;
  d = sys[unit_sys[u]];
  j = (int)u - d.unit0;
  return true;
}

// Fresh systems: x = 0, r = b, z = D^-1 b, p = z; partials r.r, r.z
__global__ void __launch_bounds__(kBlock)
    grp_init_kernel(const SysDesc *__restrict__ sys, const int *__restrict__ unit_sys, int U) {
  SysDesc d;
  int j;
  if (!unit_of(sys, unit_sys, U, d, j)) return;
  if (d.st->status != kFresh) return;
  const long long i = (long long)j * kBlock + threadIdx.x;
  double rr = 0, rz = 0;
  if (i < d.nrow) {
    const double ri = d.b[i];
    const double zi = ri * d.invd[i];
    d.x[i] = 0.0;
    d.r[i] = ri;
    d.z[i] = zi;
    d.p[i] = zi;
    rr = ri * ri;
    rz = ri * zi;
  }
  unit_partial2(rr, rz, d.part, j);
}

__global__ void __launch_bounds__(kBlock) grp_init_final(const SysDesc *__restrict__ sys) {
  const SysDesc d = sys[blockIdx.x];
  CgState<double> *st = d.st;
  if (st->status != kFresh) return;
  const double rr = units_sum(d.part, d.nunits, 2, 0);
  const double rz = units_sum(d.part, d.nunits, 2, 1);
  if (threadIdx.x == 0) {
    const double res0 = sqrt(rr);
    st->rr = rr;
    st->rz = rz;
    st->res0 = res0;
    st->res = res0;
    st->iter = 0;
    st->status = check_state(0, res0, res0, st->maxit, st->atol, st->rtol, st->dtol);
  }
}

// K1: Ap = A p, partial p.Ap per unit
__global__ void __launch_bounds__(kBlock)
    grp_spmv_kernel(const SysDesc *__restrict__ sys, const int *__restrict__ unit_sys, int U) {
  SysDesc d;
  int j;
  if (!unit_of(sys, unit_sys, U, d, j)) return;
  if (d.st->status != kIterate) return;
  const SellOp<double, int> op = sys_op(d);
  TileRow tr;
  const double s = op.row_sum(j, d.p, 0.0, tr, nullptr);
  double part = 0;
  if (tr.has_row) {
    d.Ap[tr.row] = s;
    part = s * d.p[tr.row];
  }
  __shared__ double lds[kBlock / kWave];
  const double t = block_sum(part, lds);
  if (threadIdx.x == 0) d.part[j] = t;
}

// F1: alpha = r.z / p.Ap, one workgroup per system
__global__ void __launch_bounds__(kBlock) grp_alpha_final(const SysDesc *__restrict__ sys) {
  const SysDesc d = sys[blockIdx.x];
  CgState<double> *st = d.st;
  if (st->status != kIterate) return;
  const double pAp = units_sum(d.part, d.nunits, 1, 0);
  if (threadIdx.x == 0) {
    st->pAp = pAp;
    st->alpha = st->rz / pAp;
    if (pAp == 0.0 || !(pAp == pAp)) st->status = kBreakdown;
  }
}

// K2: x += alpha p; r -= alpha Ap; z = D^-1 r; partials r.r, r.z
__global__ void __launch_bounds__(kBlock)
    grp_update_kernel(const SysDesc *__restrict__ sys, const int *__restrict__ unit_sys, int U) {
  SysDesc d;
  int j;
  if (!unit_of(sys, unit_sys, U, d, j)) return;
  if (d.st->status != kIterate) return;
  const double alpha = d.st->alpha, malpha = -alpha;
  const long long i = (long long)j * kBlock + threadIdx.x;
  double rr = 0, rz = 0;
  if (i < d.nrow) {
    d.x[i] += alpha * d.p[i];
    const double ri = d.r[i] + malpha * d.Ap[i];
    const double zi = ri * d.invd[i];
    d.r[i] = ri;
    d.z[i] = zi;
    rr = ri * ri;
    rz = ri * zi;
  }
  unit_partial2(rr, rz, d.part, j);
}

// F2: res, stopping test, beta; one workgroup per system
__global__ void __launch_bounds__(kBlock) grp_beta_final(const SysDesc *__restrict__ sys) {
  const SysDesc d = sys[blockIdx.x];
  CgState<double> *st = d.st;
  if (st->status != kIterate) return;
  const double rr = units_sum(d.part, d.nunits, 2, 0);
  const double rz = units_sum(d.part, d.nunits, 2, 1);
  if (threadIdx.x == 0) {
    const int iter = st->iter + 1;
    const double res = sqrt(rr);
    st->iter = iter;
    st->rr = rr;
    st->res = res;
    const int s = check_state(iter, res, st->res0, st->maxit, st->atol, st->rtol, st->dtol);
    if (s == kIterate) {
      st->beta = rz / st->rz;
      st->rz = rz;
    }
    st->status = s;
  }
}

// K3: p = beta p + z
__global__ void __launch_bounds__(kBlock)
    grp_p_kernel(const SysDesc *__restrict__ sys, const int *__restrict__ unit_sys, int U) {
  SysDesc d;
  int j;
  if (!unit_of(sys, unit_sys, U, d, j)) return;
  if (d.st->status != kIterate) return;
  const double beta = d.st->beta;
  const long long i = (long long)j * kBlock + threadIdx.x;
  if (i < d.nrow) d.p[i] = beta * d.p[i] + d.z[i];
}

// CSR -> SELL-64 of one system (thread per row), and its inverse diagonal
// (1/a_ii formed in double as extract_invdiagelements does,
// lmatrix_csr_cpu.cc:2419-2442; 0 where the diagonal is missing or zero,
// counted in *bad).
__global__ void __launch_bounds__(kBlock)
    sys_fill_kernel(int nrow, const int *__restrict__ rp, const int *__restrict__ ci,
                    const double *__restrict__ val, const long long *__restrict__ soff,
                    int *__restrict__ col, double *__restrict__ sval, double *__restrict__ invd,
                    unsigned int *bad) {
  constexpr int P = SellChunk<double>::P;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < nrow; i += gridDim.x * kBlock) {
    const int s = i / kSellRows, lane = i % kSellRows;
    const long long base = soff[s] + (long long)P * lane;
    const int k0 = rp[i], len = rp[i + 1] - k0;
    double dg = 0;
    for (int j = 0; j < len; ++j) {
      const long long pos = base + (long long)kSellRows * P * (j / P) + (j % P);
      const int c = ci[k0 + j];
      const double v = val[k0 + j];
      col[pos] = c;
      sval[pos] = v;
      if (c == i && dg == 0.0) dg = v;
    }
    if (dg == 0.0) {
      invd[i] = 0.0;
      atomicAdd(bad, 1u);
    } else {
      invd[i] = 1.0 / dg;
    }
  }
}

// ---- host side ------------------------------------------------------------------

// First-fit allocator over the arena with coalescing of free neighbours.
class Arena {
 public:
  void reset(size_t size) {
    free_.clear();
    if (size) free_[0] = size;
  }
  bool alloc(size_t bytes, size_t *off) {
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= bytes) {
        *off = it->first;
        const size_t rest = it->second - bytes;
        const size_t start = it->first + bytes;
        free_.erase(it);
        if (rest) free_[start] = rest;
        return true;
      }
    }
    return false;
  }
  void release(size_t off, size_t bytes) {
    auto it = free_.emplace(off, bytes).first;
    auto next = std::next(it);
    if (next != free_.end() && it->first + it->second == next->first) {
      it->second += next->second;
      free_.erase(next);
    }
    if (it != free_.begin()) {
      auto prev = std::prev(it);
      if (prev->first + prev->second == it->first) {
        prev->second += it->second;
        free_.erase(it);
      }
    }
  }
  size_t largest() const {
    size_t m = 0;
    for (const auto &f : free_) m = std::max(m, f.second);
    return m;
  }
  size_t total_free() const {
    size_t m = 0;
    for (const auto &f : free_) m += f.second;
    return m;
  }

 private:
  std::map<size_t, size_t> free_;  // offset -> size
};

enum JobState { kQueued = 0, kActive = 1, kDone = 2 };

struct Job {
  long long ticket = 0;
  int nrow = 0, nnz = 0, nslice = 0, nunits = 0;
  const int *rp = nullptr, *ci = nullptr;
  const double *val = nullptr, *b = nullptr;
  double *x_out = nullptr;
  double atol = 0, rtol = 0;
  int maxit = 0;
  std::vector<long long> soff;
  std::vector<int> slen;
  std::vector<unsigned char> rlen;
  long long entries = 0;
  size_t bytes = 0, off = 0;
  int slot = -1;
  int state = kQueued;
  SysDesc d{};
  // device-side block layout (offsets inside the job's arena block)
  size_t o_rp = 0, o_ci = 0, o_val = 0, o_soff = 0, o_slen = 0, o_rlen = 0, o_col = 0,
         o_sval = 0, o_invd = 0, o_b = 0, o_x = 0, o_r = 0, o_p = 0, o_z = 0, o_Ap = 0,
         o_part = 0;
  std::chrono::steady_clock::time_point t_submit, t_admit, t_done;
  int iters = 0, status = 0, bad_diag = 0;
  double res_init = 0, res_final = 0;
};

// Lays out the job's arrays and computes its footprint (256-byte aligned).
void plan_job(Job &jb) {
  constexpr int P = SellChunk<double>::P;
  jb.nslice = (jb.nrow + kSellRows - 1) / kSellRows;
  jb.nunits = (jb.nrow + kBlock - 1) / kBlock;
  jb.slen.assign(jb.nslice, 0);
  jb.rlen.assign(jb.nrow, 0);
  jb.soff.assign((size_t)jb.nslice + 1, 0);
  long long tot = 0;
  for (int s = 0; s < jb.nslice; ++s) {
    int m = 0;
    for (int l = 0; l < kSellRows; ++l) {
      const int r = s * kSellRows + l;
      if (r >= jb.nrow) break;
      const int len = jb.rp[r + 1] - jb.rp[r];
      jb.rlen[r] = (unsigned char)len;
      m = std::max(m, len);
    }
    jb.slen[s] = (m + P - 1) / P * P;
    jb.soff[s] = tot;
    tot += (long long)kSellRows * jb.slen[s];
  }
  jb.soff[jb.nslice] = tot;
  jb.entries = tot;
  size_t o = 0;
  auto take = [&](size_t bytes) {
    const size_t at = o;
    o = align256(o + bytes);
    return at;
  };
  const size_t n = (size_t)jb.nrow;
  jb.o_rp = take(4 * (n + 1));
  jb.o_ci = take(4 * (size_t)jb.nnz);
  jb.o_val = take(8 * (size_t)jb.nnz);
  jb.o_soff = take(8 * ((size_t)jb.nslice + 1));
  jb.o_slen = take(4 * (size_t)jb.nslice);
  jb.o_rlen = take(n);
  jb.o_col = take(4 * (size_t)tot);
  jb.o_sval = take(8 * (size_t)tot);
  jb.o_invd = take(8 * n);
  jb.o_b = take(8 * n);
  jb.o_x = take(8 * n);
  jb.o_r = take(8 * n);
  jb.o_p = take(8 * n);
  jb.o_z = take(8 * n);
  jb.o_Ap = take(8 * n);
  jb.o_part = take(16 * (size_t)jb.nunits);
  jb.bytes = o;
}

}  // namespace

struct hflmp_server {
  hflmp_ctx *ctx = nullptr;
  size_t budget = 0;
  char *arena = nullptr;
  Arena alloc;
  int max_active = 0, chunk = 0, use_graph = 1;
  CgState<double> *d_states = nullptr;  // [max_active]
  CgState<double> *h_states = nullptr;  // pinned mirror
  SysDesc *d_sys = nullptr;             // [max_active]
  int *d_unit_sys = nullptr;
  size_t unit_cap = 0;
  unsigned int *d_bad = nullptr;
  std::vector<SysDesc> h_sys;
  std::vector<int> h_unit_sys;
  std::vector<long long> slot_job;  // slot -> ticket, -1 free
  std::vector<int> order;           // active slots in table order
  std::deque<long long> queue;
  std::map<long long, Job> jobs;
  long long next_ticket = 1;
  bool set_changed = true;
  int U = 0, S = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  long long steps = 0, graph_captures = 0;
};

namespace {

using Clock = std::chrono::steady_clock;

void drop_graph(hflmp_server *sv) {
  if (sv->graph_exec) (void)hipGraphExecDestroy(sv->graph_exec);
  if (sv->graph) (void)hipGraphDestroy(sv->graph);
  sv->graph_exec = nullptr;
  sv->graph = nullptr;
}

// The work of one step on the stream: initial residuals of fresh systems,
// then `chunk` iterations.
void enqueue_chunk(hflmp_server *sv, hipStream_t s) {
  const unsigned gU = (unsigned)(((sv->U + kXcds - 1) / kXcds) * kXcds);
  const unsigned gS = (unsigned)sv->S;
  grp_init_kernel<<<gU, kBlock, 0, s>>>(sv->d_sys, sv->d_unit_sys, sv->U);
  grp_init_final<<<gS, kBlock, 0, s>>>(sv->d_sys);
  for (int it = 0; it < sv->chunk; ++it) {
    grp_spmv_kernel<<<gU, kBlock, 0, s>>>(sv->d_sys, sv->d_unit_sys, sv->U);
    grp_alpha_final<<<gS, kBlock, 0, s>>>(sv->d_sys);
    grp_update_kernel<<<gU, kBlock, 0, s>>>(sv->d_sys, sv->d_unit_sys, sv->U);
    grp_beta_final<<<gS, kBlock, 0, s>>>(sv->d_sys);
    grp_p_kernel<<<gU, kBlock, 0, s>>>(sv->d_sys, sv->d_unit_sys, sv->U);
  }
}

int rebuild_tables(hflmp_server *sv) {
  hflmp_ctx *ctx = sv->ctx;
  sv->h_sys.clear();
  sv->h_unit_sys.clear();
  int unit0 = 0;
  for (int slot : sv->order) {
    Job &jb = sv->jobs[sv->slot_job[slot]];
    jb.d.unit0 = unit0;
    sv->h_sys.push_back(jb.d);
    for (int u = 0; u < jb.nunits; ++u) sv->h_unit_sys.push_back((int)sv->h_sys.size() - 1);
    unit0 += jb.nunits;
  }
  sv->S = (int)sv->h_sys.size();
  sv->U = unit0;
  if ((size_t)sv->U > sv->unit_cap) {
    if (sv->d_unit_sys) (void)hipFree(sv->d_unit_sys);
    sv->unit_cap = std::max((size_t)sv->U, 2 * sv->unit_cap);
    if (hipMalloc(&sv->d_unit_sys, sizeof(int) * sv->unit_cap) != hipSuccess)
      return set_error(ctx, HFLMP_ENOMEM, "server: unit table allocation failed");
  }
  if (sv->S) {
    HFLMP_HIP_CHECK(ctx, hipMemcpyAsync(sv->d_sys, sv->h_sys.data(), sizeof(SysDesc) * sv->S,
                                        hipMemcpyHostToDevice, ctx->stream));
    HFLMP_HIP_CHECK(ctx, hipMemcpyAsync(sv->d_unit_sys, sv->h_unit_sys.data(),
                                        sizeof(int) * sv->U, hipMemcpyHostToDevice,
                                        ctx->stream));
  }
  drop_graph(sv);
  sv->set_changed = false;
  return HFLMP_OK;
}

int admit(hflmp_server *sv, Job &jb, int slot) {
  hflmp_ctx *ctx = sv->ctx;
  hipStream_t s = ctx->stream;
  char *base = sv->arena + jb.off;
  auto dp = [&](size_t o) { return base + o; };
  const size_t n = (size_t)jb.nrow;
  HFLMP_HIP_CHECK(ctx, hipMemcpyAsync(dp(jb.o_rp), jb.rp, 4 * (n + 1), hipMemcpyHostToDevice, s));
  if (jb.nnz) {
    HFLMP_HIP_CHECK(ctx, hipMemcpyAsync(dp(jb.o_ci), jb.ci, 4 * (size_t)jb.nnz,
                                        hipMemcpyHostToDevice, s));
    HFLMP_HIP_CHECK(ctx, hipMemcpyAsync(dp(jb.o_val), jb.val, 8 * (size_t)jb.nnz,
                                        hipMemcpyHostToDevice, s));
  }
  HFLMP_HIP_CHECK(ctx, hipMemcpyAsync(dp(jb.o_soff), jb.soff.data(), 8 * jb.soff.size(),
                                      hipMemcpyHostToDevice, s));
  HFLMP_HIP_CHECK(ctx, hipMemcpyAsync(dp(jb.o_slen), jb.slen.data(), 4 * jb.slen.size(),
                                      hipMemcpyHostToDevice, s));
  HFLMP_HIP_CHECK(ctx, hipMemcpyAsync(dp(jb.o_rlen), jb.rlen.data(), n, hipMemcpyHostToDevice, s));
  HFLMP_HIP_CHECK(ctx, hipMemcpyAsync(dp(jb.o_b), jb.b, 8 * n, hipMemcpyHostToDevice, s));
  int rc = hflmp_fill_zero(ctx, dp(jb.o_col), 4 * (size_t)jb.entries);
  if (!rc) rc = hflmp_fill_zero(ctx, dp(jb.o_sval), 8 * (size_t)jb.entries);
  if (rc) return rc;
  sys_fill_kernel<<<grid_for(jb.nrow, 1), kBlock, 0, s>>>(
      jb.nrow, reinterpret_cast<const int *>(dp(jb.o_rp)),
      reinterpret_cast<const int *>(dp(jb.o_ci)), reinterpret_cast<const double *>(dp(jb.o_val)),
      reinterpret_cast<const long long *>(dp(jb.o_soff)), reinterpret_cast<int *>(dp(jb.o_col)),
      reinterpret_cast<double *>(dp(jb.o_sval)), reinterpret_cast<double *>(dp(jb.o_invd)),
      sv->d_bad + slot);
  HFLMP_LAUNCH_CHECK(ctx);
  CgState<double> st{};
  st.status = kFresh;
  st.maxit = jb.maxit;
  st.atol = jb.atol;
  st.rtol = jb.rtol;
  st.dtol = 1e300;
  // the pinned mirror slot is free (the slot was retired) and is only read
  // after the next state copy
  sv->h_states[slot] = st;
  HFLMP_HIP_CHECK(ctx, hipMemcpyAsync(sv->d_states + slot, sv->h_states + slot, sizeof(st),
                                      hipMemcpyHostToDevice, s));
  SysDesc &d = jb.d;
  d.nrow = jb.nrow;
  d.nslice = jb.nslice;
  d.nunits = jb.nunits;
  d.soff = reinterpret_cast<const long long *>(dp(jb.o_soff));
  d.slen = reinterpret_cast<const int *>(dp(jb.o_slen));
  d.rlen = reinterpret_cast<const unsigned char *>(dp(jb.o_rlen));
  d.col = reinterpret_cast<const int *>(dp(jb.o_col));
  d.sval = reinterpret_cast<const double *>(dp(jb.o_sval));
  d.invd = reinterpret_cast<const double *>(dp(jb.o_invd));
  d.b = reinterpret_cast<const double *>(dp(jb.o_b));
  d.x = reinterpret_cast<double *>(dp(jb.o_x));
  d.r = reinterpret_cast<double *>(dp(jb.o_r));
  d.p = reinterpret_cast<double *>(dp(jb.o_p));
  d.z = reinterpret_cast<double *>(dp(jb.o_z));
  d.Ap = reinterpret_cast<double *>(dp(jb.o_Ap));
  d.part = reinterpret_cast<double *>(dp(jb.o_part));
  d.st = sv->d_states + slot;
  jb.slot = slot;
  jb.state = kActive;
  jb.t_admit = Clock::now();
  return HFLMP_OK;
}

}  // namespace

extern "C" {

int hflmp_server_create(hflmp_ctx *ctx, const hflmp_server_config *cfg, hflmp_server **out) {
  if (!ctx) return HFLMP_EINVAL;
  HFLMP_REQUIRE(ctx, cfg && out, "server_create: null pointer");
  *out = nullptr;
  HFLMP_REQUIRE(ctx, cfg->hbm_budget >= (1u << 20) && cfg->max_active >= 1 &&
                         cfg->max_active <= 4096 && cfg->chunk_iters >= 1 &&
                         cfg->chunk_iters <= 256,
                "server_create: need hbm_budget >= 1 MiB, 1 <= max_active <= 4096, "
                "1 <= chunk_iters <= 256");
  auto *sv = new hflmp_server();
  sv->ctx = ctx;
  sv->budget = cfg->hbm_budget;
  sv->max_active = cfg->max_active;
  sv->chunk = cfg->chunk_iters;
  sv->use_graph = cfg->use_graphs ? 1 : 0;
  auto fail = [&](const char *what) {
    hflmp_server_destroy(sv);
    return set_error(ctx, HFLMP_ENOMEM, "server_create: %s", what);
  };
  if (hipMalloc(&sv->arena, sv->budget) != hipSuccess) return fail("arena allocation failed");
  if (hipMalloc(&sv->d_states, sizeof(CgState<double>) * sv->max_active) != hipSuccess ||
      hipMalloc(&sv->d_sys, sizeof(SysDesc) * sv->max_active) != hipSuccess ||
      hipMalloc(&sv->d_bad, sizeof(unsigned int) * sv->max_active) != hipSuccess ||
      hipHostMalloc(&sv->h_states, sizeof(CgState<double>) * sv->max_active) != hipSuccess)
    return fail("table allocation failed");
  sv->alloc.reset(sv->budget);
  sv->slot_job.assign(sv->max_active, -1);
  if (hflmp_fill_zero(ctx, sv->d_bad, sizeof(unsigned int) * sv->max_active))
    return fail("fill");
  *out = sv;
  return HFLMP_OK;
}

void hflmp_server_destroy(hflmp_server *sv) {
  if (!sv) return;
  if (sv->ctx) (void)hipStreamSynchronize(sv->ctx->stream);
  drop_graph(sv);
  (void)hipFree(sv->arena);
  (void)hipFree(sv->d_states);
  (void)hipFree(sv->d_sys);
  (void)hipFree(sv->d_unit_sys);
  (void)hipFree(sv->d_bad);
  if (sv->h_states) (void)hipHostFree(sv->h_states);
  delete sv;
}

size_t hflmp_server_footprint(int nrow, int nnz, const int *row_ptr) {
  if (nrow <= 0 || nnz < 0 || !row_ptr) return 0;
  Job jb;
  jb.nrow = nrow;
  jb.nnz = nnz;
  jb.rp = row_ptr;
  plan_job(jb);
  return jb.bytes;
}

int hflmp_server_submit(hflmp_server *sv, int nrow, int nnz, const int *row_ptr,
                        const int *col, const double *val, const double *b, double *x,
                        double atol, double rtol, int maxit, long long *ticket) {
  if (!sv) return HFLMP_EINVAL;
  hflmp_ctx *ctx = sv->ctx;
  HFLMP_REQUIRE(ctx, nrow > 0 && nnz >= 0 && row_ptr && b && x && ticket &&
                         (nnz == 0 || (col && val)),
                "server_submit: bad system");
  HFLMP_REQUIRE(ctx, maxit > 0 && atol >= 0 && rtol >= 0, "server_submit: bad tolerances");
  HFLMP_REQUIRE(ctx, row_ptr[0] == 0 && row_ptr[nrow] == nnz,
                "server_submit: row_ptr does not span nnz");
  for (int r = 0; r < nrow; ++r)
    HFLMP_REQUIRE(ctx, row_ptr[r + 1] - row_ptr[r] >= 0 && row_ptr[r + 1] - row_ptr[r] <= 255,
                  "server_submit: row %d has %d entries (0..255 supported)", r,
                  row_ptr[r + 1] - row_ptr[r]);
  Job jb;
  jb.ticket = sv->next_ticket++;
  jb.nrow = nrow;
  jb.nnz = nnz;
  jb.rp = row_ptr;
  jb.ci = col;
  jb.val = val;
  jb.b = b;
  jb.x_out = x;
  jb.atol = atol;
  jb.rtol = rtol;
  jb.maxit = maxit;
  plan_job(jb);
  HFLMP_REQUIRE(ctx, jb.bytes <= sv->budget,
                "server_submit: footprint %zu B exceeds the HBM budget %zu B", jb.bytes,
                sv->budget);
  jb.t_submit = Clock::now();
  *ticket = jb.ticket;
  sv->queue.push_back(jb.ticket);
  sv->jobs.emplace(jb.ticket, std::move(jb));
  return HFLMP_OK;
}

int hflmp_server_step(hflmp_server *sv, int *n_done) {
  if (!sv) return HFLMP_EINVAL;
  hflmp_ctx *ctx = sv->ctx;
  hipStream_t s = ctx->stream;
  int done = 0;
  // admit first come, first served while a slot is free and the block fits
  while (!sv->queue.empty()) {
    int slot = -1;
    for (int k = 0; k < sv->max_active; ++k)
      if (sv->slot_job[k] < 0) {
        slot = k;
        break;
      }
    if (slot < 0) break;
    Job &jb = sv->jobs[sv->queue.front()];
    if (!sv->alloc.alloc(jb.bytes, &jb.off)) break;
    int rc = admit(sv, jb, slot);
    if (rc) return rc;
    sv->slot_job[slot] = jb.ticket;
    sv->order.push_back(slot);
    sv->queue.pop_front();
    sv->set_changed = true;
  }
  if (sv->set_changed) {
    int rc = rebuild_tables(sv);
    if (rc) return rc;
  }
  if (sv->S > 0) {
    if (sv->use_graph) {
      if (!sv->graph_exec) {
        HFLMP_HIP_CHECK(ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        enqueue_chunk(sv, s);
        HFLMP_HIP_CHECK(ctx, hipStreamEndCapture(s, &sv->graph));
        HFLMP_HIP_CHECK(ctx, hipGraphInstantiate(&sv->graph_exec, sv->graph, nullptr, nullptr, 0));
        ++sv->graph_captures;
      }
      HFLMP_HIP_CHECK(ctx, hipGraphLaunch(sv->graph_exec, s));
    } else {
      enqueue_chunk(sv, s);
      HFLMP_LAUNCH_CHECK(ctx);
    }
    ++sv->steps;
    HFLMP_HIP_CHECK(ctx, hipMemcpyAsync(sv->h_states, sv->d_states,
                                        sizeof(CgState<double>) * sv->max_active,
                                        hipMemcpyDeviceToHost, s));
    HFLMP_HIP_CHECK(ctx, hipStreamSynchronize(s));
    // retire the systems that stopped
    std::vector<int> keep;
    for (int slot : sv->order) {
      const CgState<double> &st = sv->h_states[slot];
      Job &jb = sv->jobs[sv->slot_job[slot]];
      if (st.status == kIterate || st.status == kFresh) {
        keep.push_back(slot);
        continue;
      }
      HFLMP_HIP_CHECK(ctx, hipMemcpy(jb.x_out, jb.d.x, 8 * (size_t)jb.nrow,
                                     hipMemcpyDeviceToHost));
      unsigned int bad = 0;
      HFLMP_HIP_CHECK(ctx, hipMemcpy(&bad, sv->d_bad + slot, sizeof(bad), hipMemcpyDeviceToHost));
      if (bad) {
        const int rc = hflmp_fill_zero(ctx, sv->d_bad + slot, sizeof(bad));
        if (rc) return rc;
      }
      jb.iters = st.iter;
      jb.status = st.status;
      jb.res_init = st.res0;
      jb.res_final = st.res;
      jb.bad_diag = (int)bad;
      jb.t_done = Clock::now();
      jb.state = kDone;
      sv->alloc.release(jb.off, jb.bytes);
      sv->slot_job[slot] = -1;
      sv->set_changed = true;
      ++done;
    }
    sv->order.swap(keep);
  }
  if (n_done) *n_done = done;
  return HFLMP_OK;
}

int hflmp_server_poll(hflmp_server *sv, long long ticket, hflmp_server_result *res) {
  if (!sv || !res) return HFLMP_EINVAL;
  auto it = sv->jobs.find(ticket);
  if (it == sv->jobs.end())
    return set_error(sv->ctx, HFLMP_EINVAL, "server_poll: unknown ticket %lld", ticket);
  const Job &jb = it->second;
  res->state = jb.state;
  res->iters = jb.iters;
  res->status = jb.state == kDone ? (jb.status == kIterate ? kMaxIts : jb.status) : -1;
  res->res_init = jb.res_init;
  res->res_final = jb.res_final;
  res->zero_diagonals = jb.bad_diag;
  res->footprint = jb.bytes;
  auto sec = [](Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double>(b - a).count();
  };
  res->queued_seconds = jb.state >= kActive ? sec(jb.t_submit, jb.t_admit) : -1.0;
  res->latency_seconds = jb.state == kDone ? sec(jb.t_submit, jb.t_done) : -1.0;
  return HFLMP_OK;
}

int hflmp_server_release(hflmp_server *sv, long long ticket) {
  if (!sv) return HFLMP_EINVAL;
  auto it = sv->jobs.find(ticket);
  if (it == sv->jobs.end() || it->second.state != kDone)
    return set_error(sv->ctx, HFLMP_EINVAL, "server_release: ticket %lld is not done", ticket);
  sv->jobs.erase(it);
  return HFLMP_OK;
}

int hflmp_server_stats(const hflmp_server *sv, hflmp_server_stats_t *out) {
  if (!sv || !out) return HFLMP_EINVAL;
  out->queued = (int)sv->queue.size();
  out->active = (int)sv->order.size();
  out->active_units = sv->U;
  out->arena_free = sv->alloc.total_free();
  out->arena_largest_free = sv->alloc.largest();
  out->steps = sv->steps;
  out->graph_captures = sv->graph_captures;
  return HFLMP_OK;
}

}  // extern "C"
