// libbiefhip: MI355X (gfx950) implementation of TELEMAC's BIEF linear-system
// path (EBE mat-vec, diagonal preconditioning, CG / BiCGSTAB / GMRES).
// C ABI in include/bief_hip.h.
//
// Design (HBM-bound integer/FP streaming, no MFMA):
//  * nodes renumbered (Hilbert) and cut into blocks of consecutive ids; one
//    workgroup per block owns its nodes (layout.cpp);
//  * a workgroup stages the vector it multiplies into LDS for its owned and
//    halo nodes, walks its elements colour by colour (no two elements of a
//    colour share an owned node) accumulating in LDS, then writes its owned
//    nodes once -- deterministic, atomic-free scatter-add;
//  * element data are structure of arrays in slot order (coalesced): 16-bit
//    local node ids and the off-diagonal coefficients;
//  * CG keeps its scalars on the device and never synchronises workgroups
//    inside a kernel (on multi-XCD MI355X an agent-scope fence is an L2
//    write-back): each kernel writes one partial sum per workgroup, and every
//    workgroup of the next kernel reduces those partials itself in the same
//    fixed order, so all of them derive identical scalars (rho, beta, the
//    stop test). Scalars written for the host are double-buffered by
//    iteration parity so no kernel reads a slot it writes. Iterations run
//    back to back without host round trips and are bitwise reproducible.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/bief_hip.h"
#include "layout.h"

namespace bhip {

constexpr int kThreads = 256;
static_assert(kThreads == kPhaseSlots, "colour phases are balanced for this workgroup size");
constexpr int kWaves = kThreads / 64;

#define HIP_OK(expr)                                              \
  do {                                                            \
    hipError_t e_ = (expr);                                       \
    if (e_ != hipSuccess) {                                       \
      fprintf(stderr, "libbiefhip: %s failed: %s (%s:%d)\n", #expr, \
              hipGetErrorString(e_), __FILE__, __LINE__);         \
      return BHIP_EDEVICE;                                        \
    }                                                             \
  } while (0)

// ---------------------------------------------------------------------------
// solver state (device resident)
// ---------------------------------------------------------------------------
enum : int { ST_RUNNING = 0, ST_CONVERGED = 1, ST_FINAL = 2, ST_MAXITER = 3 };

// Iteration m of a CG solve runs U_m (vector update, parity p = m & 1) then
// E_m (direction + EBE mat-vec). Slot [p] is written by iteration m and read
// by iteration m + 1.
struct SolveState {
  double xl;        // max(||b'||^2, 1) (gracjg.F:189-209)
  double eps2xl;    // stop threshold on ||r||^2 (inner solves: relative)
  double rr0;       // ||r_0||^2 of the solve (inner solves scale eps2xl by it)
  double dot;       // result of k_reduce_into
  double rr[2];     // RMRM of iteration m, by parity
  double rho[2];    // RO of iteration m (informational)
  int m[2];         // iteration number M, by parity
  int status[2];    // ST_RUNNING or ST_FINAL (NITMAX reached), by parity
  int done;         // sticky: 0, ST_CONVERGED or ST_MAXITER
  int nitmax;
  int final_m;      // M at exit
  int use_rr0;      // threshold = eps2xl * rr0
  double final_rr;  // RMRM at exit
  int final_flag;   // set by the update kernel that applied the last X update
  int zero_rhs;     // ||b'|| < 1e-15 (gracjg.F:194-202)
  double xl_raw;    // ||b'||^2
  // mixed precision outer loop (device-side control)
  double eps2_outer;   // outer stop threshold on the FP64 ||r'||^2
  double rr_outer;     // last FP64 ||r'||^2
  double inner_eps2;   // inner relative threshold
  int inner_nitmax;    // iteration cap of one inner solve
  int nitmax_total;    // NITMAX of the whole mixed solve (inner iterations)
  int outer_done;      // the outer loop has stopped (outer_status says why)
  int outer_status;    // ST_CONVERGED or ST_MAXITER once outer_done
  int outer_count;     // inner solves started
  int inner_total;     // inner iterations of finished inner solves
  int done1;           // inner solve finished: (M << 2) | status, set by a kernel
                       // after the one that applied the last update
  int fresh;           // an outer test started an inner solve whose first
                       // kernel has not run yet (set by k_outer_check,
                       // cleared by the first update kernel of that solve)
  int xpar;            // which of the two FP64 solution buffers holds x
  int dpar;            // reliable updates: parity of the direction the next inner
                       // start continues from, -1 for a fresh start
  int reliable;        // mixed path: inner CG continued across outer steps
  int inner_pairs[8];  // CG iterations (update + direction launches) of the
                       // first inner solves, the next solve's launch plan
  int stop;            // CG kernels of this solve return at once. Set by the
                       // kernel after the one that set done, so that no
                       // workgroup of a direction kernel skips its work on a
                       // flag another workgroup of the same launch has set
  long long t_start;   // device clock at the set-up of the solve
  long long t_done;    // device clock when this system stopped
};

// ---------------------------------------------------------------------------
// deterministic reductions
// ---------------------------------------------------------------------------
__device__ inline double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// fixed-order block sum; result valid in thread 0
__device__ inline double block_sum(double v, double* sh) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += sh[i];
  }
  __syncthreads();
  return s;
}

// Fixed-order sum of n partials computed by every workgroup that calls it
// (identical result in all of them); valid in all threads.
__device__ inline double reduce_partials(const double* __restrict__ partial, int n,
                                         double* sh, double* bcast) {
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) s += partial[i];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) *bcast = s;
  __syncthreads();
  s = *bcast;
  __syncthreads();
  return s;
}

// reduce_partials in two steps, for kernels that want other loads in
// flight meanwhile: load_partials issues this thread's share of the partial
// sums first (indices clamped, no branches around the loads); sum_partials
// adds them in reduce_partials' order, so the result is the same to the bit.
template <int P>
__device__ inline void load_partials(double (&pv)[P], const double* __restrict__ partial,
                                     int n) {
#pragma unroll
  for (int k = 0; k < P; ++k) pv[k] = partial[min((int)threadIdx.x + k * kThreads, n - 1)];
}
template <int P>
__device__ inline double sum_partials(const double (&pv)[P], const double* __restrict__ partial,
                                      int n, double* sh, double* bcast) {
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < P; ++k)
    if ((int)threadIdx.x + k * kThreads < n) s += pv[k];
  for (int i = threadIdx.x + P * kThreads; i < n; i += blockDim.x) s += partial[i];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) *bcast = s;
  __syncthreads();
  s = *bcast;
  __syncthreads();
  return s;
}

// XCD-aware block order: workgroup g is dispatched to XCD g % 8; each XCD
// gets a contiguous range of node blocks so that blocks sharing halo nodes
// share an L2. Bijective for any block count.
__device__ inline int xcd_block(int g, int nb) {
  const int q = nb >> 3, r = nb & 7;
  const int xcd = g & 7, k = g >> 3;
  return xcd * q + (xcd < r ? xcd : r) + k;
}

__device__ inline int load_relaxed(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void store_relaxed(int* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------
// EBE kernel
// ---------------------------------------------------------------------------
enum : int { DIAG_ZERO = 0, DIAG_ONE = 1, DIAG_Q = 2 };
// how the multiplied vector is obtained for each local node
enum : int { STAGE_PLAIN = 0, STAGE_CGDIR = 1, STAGE_XUPD = 2 };
// what happens at owned nodes after the element loop
enum : int { EPI_OP = 0, EPI_RESID = 1, EPI_CGAD = 2, EPI_RESID_MIXED = 3, EPI_DOTW = 4,
             EPI_DOT2 = 5 };
// whether the kernel writes a partial sum per workgroup / runs the CG control
enum : int { FIN_NONE = 0, FIN_PARTIAL = 1, FIN_CG = 2 };

template <typename T, typename TA>
struct EbeArgs {
  // layout
  const int32_t* node_start;
  const int32_t* halo_start;
  const int32_t* halo;
  const int32_t* color_off;
  const uint16_t* lid;
  int col_ids;                // prisms numbered by column: ids 3..5 = ids 0..2 + 1
  int nslot_pad;
  int nblocks;
  // matrix
  const TA* xa;   // [ncoef][nslot_pad]
  const TA* da;   // [npoin] (DIAG_Q)
  // stage inputs
  const T* y;                 // PLAIN: multiplied vector
  const T* r;                 // CGDIR: residual
  const T* dold;              // CGDIR: previous direction
  T* dnew;                    // CGDIR: new direction (owned nodes written)
  T* xzero;                   // CGDIR (first kernel of an inner solve): its
                              // solution, zeroed at owned nodes
  double* xb0;                // XUPD: the two FP64 solution buffers; the
  double* xb1;                // system's xpar names the current one, x - delta
                              // is written to the other (owned nodes)
  const float* delta;         // XUPD: inner correction
  // epilogue
  T* out;
  const T* xin;
  double c;
  int op;
  const double* b;            // RESID: right-hand side (FP64)
  const T* w;                 // DOTW: out = A y, partial w.out
  double* partial2;           // DOT2: out = A y, partials out.y and out.out
  float* r32;                 // RESID_MIXED: inner residual -(A x - b)
  // reductions / state
  double* partial;            // this kernel's partials (one per workgroup)
  const double* part_in;      // FIN_CG: partials of the preceding update kernel
  int n_part_in;
  SolveState* st;
  int parity;                 // FIN_CG: parity of the iteration
  int skip_unless_fresh;      // first direction kernel of an inner solve: runs
                              // only for systems whose outer test started one
  int skip_if_done;           // kernels of a running solve after it finished
  int skip_if_outer_done;     // mixed outer residual: skip systems that stopped
                              // or whose inner solve is still running
  // ensembles: system of each block and block range of each system (NULL
  // for one system: all blocks, partials [0, n_part_in))
  const int32_t* block_sys;
  const int32_t* sys_blocks;
};

// column of the coefficient array holding A(i, j), i != j
template <int NDP, bool SYMQ>
__host__ __device__ constexpr int coef_col(int i, int j) {
  constexpr int npair = NDP * (NDP - 1) / 2;
  int a = i < j ? i : j, b = i < j ? j : i;
  // row-major upper pair index
  int idx = a * NDP - a * (a + 1) / 2 + (b - a - 1);
  return (i < j || !SYMQ) ? idx : npair + idx;
}

// element contributions W_k = sum_{j != k} A(k, j) Y(i_j), j increasing
// (mv0303.f:123-130, mv0606.f:117-200), scattered into the owned nodes' LDS
// accumulators
template <typename T, int NDP, bool SYMQ, bool TRANS>
__device__ inline void slot_apply(const int (&l)[NDP], const T* cf, const T* ys, T* acc,
                                  int nown) {
  T yv[NDP];
#pragma unroll
  for (int k = 0; k < NDP; ++k) yv[k] = ys[l[k]];
#pragma unroll
  for (int k = 0; k < NDP; ++k) {
    T w = (T)0;
    bool first = true;
#pragma unroll
    for (int j = 0; j < NDP; ++j) {
      if (j == k) continue;
      const int q = TRANS ? coef_col<NDP, SYMQ>(j, k) : coef_col<NDP, SYMQ>(k, j);
      const T t = cf[q] * yv[j];
      w = first ? t : w + t;
      first = false;
    }
    if (l[k] < nown) acc[l[k]] += w;
  }
}

// Element data of one colour phase for one thread (its first slot of the
// phase; phases larger than the workgroup are finished by a loop).
// local node ids of slot s; column-numbered prisms store all six (other
// kernels read them) but the product reads the bottom three only
template <int NDP>
__device__ inline void load_ids(int (&l)[NDP], const uint16_t* __restrict__ lid, size_t stride,
                                int s, int col_ids) {
  if constexpr (NDP == 6) {
    if (col_ids) {
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        l[k] = lid[(size_t)k * stride + s];
        l[k + 3] = l[k] + 1;
      }
      return;
    }
  }
#pragma unroll
  for (int k = 0; k < NDP; ++k) l[k] = lid[(size_t)k * stride + s];
}

template <typename T, int NDP, int NC>
struct PhaseSlot {
  int l[NDP];
  T cf[NC];
  bool valid;
};

template <typename T, typename TA, int NDP, int NC>
__device__ inline void load_phase(PhaseSlot<T, NDP, NC>& ps, int c, const int* sh_co, int se,
                                  const uint16_t* __restrict__ lid,
                                  const TA* __restrict__ xa, int nslot_pad, int col_ids) {
  const int s = (c < kMaxColors) ? sh_co[c] + (int)threadIdx.x : se;
  ps.valid = (c < kMaxColors) && s < sh_co[c + 1] && s < se;
  if (ps.valid) {
    load_ids<NDP>(ps.l, lid, nslot_pad, s, col_ids);
#pragma unroll
    for (int q = 0; q < NC; ++q) ps.cf[q] = (T)xa[(size_t)q * nslot_pad + s];
  }
}

// node values read by the CG staging passes (global index, R, D, AD, X)
template <typename T>
struct NodeIn {
  int g;
  T r, d;
};
// CG staging: the first node of each thread is loaded ahead of the scalar
// reduction; the others kStageUnroll at a time (their loads issued together)
#ifndef BHIP_STAGE_UNROLL
#define BHIP_STAGE_UNROLL 4
#endif
constexpr int kStagePre = 1;
#ifndef BHIP_STAGE_PRE_F32
#define BHIP_STAGE_PRE_F32 1
#endif
constexpr int kStagePreF32 = BHIP_STAGE_PRE_F32;
constexpr int kEbePart = 8;  // partial sums per thread loaded ahead in the CG kernels
constexpr int kStageUnroll = BHIP_STAGE_UNROLL;

// PF > 0: the colour phases are software pipelined: each thread keeps the
// element data of the next PF phases in flight in registers, so a phase
// waits on loads issued PF phases earlier instead of paying a full HBM round
// trip behind every barrier. PF = 0: loads inside each phase.
template <typename T, typename TA, int NDP, bool SYMQ, bool TRANS, int DIAG, int STAGE,
          int EPI, int FIN, int PF>
__global__ void __launch_bounds__(kThreads)
k_ebe(EbeArgs<T, TA> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ double sh_red[kWaves];
  __shared__ double sh_b;
  __shared__ int sh_co[kMaxColors + 1];
  constexpr int npair = NDP * (NDP - 1) / 2;
  constexpr int ncoef = SYMQ ? 2 * npair : npair;

  // XCD-aware mapping: workgroup g is dispatched to XCD g % 8; give each XCD
  // a contiguous range of node blocks so blocks sharing halo nodes share an
  // L2. Bijective for any block count.
  const int b = xcd_block(blockIdx.x, a.nblocks);
  // the system (mesh of an ensemble) of this block, its state and the range
  // of the partial sums (one per block) that belong to it
  const int sys = a.block_sys ? a.block_sys[b] : 0;
  const int pb0 = a.sys_blocks ? a.sys_blocks[sys] : 0;
  const int pb1 = a.sys_blocks ? a.sys_blocks[sys + 1] : a.n_part_in;
  if (a.skip_if_outer_done && outer_blocked(a.st + sys)) return;
  if (a.skip_if_done && load_relaxed(&a.st->done)) return;
  const int32_t* co = a.color_off + (size_t)b * (kMaxColors + 1);
  if (threadIdx.x <= kMaxColors) sh_co[threadIdx.x] = co[threadIdx.x];
  __syncthreads();
  const int se = sh_co[kMaxColors];

  // the partial sums of the previous kernel first (FIN_CG): loads issued
  // before the element and node loads below, so that adding them up waits
  // for them alone while the rest stays in flight
  double pvr[FIN == FIN_CG ? kEbePart : 1];
  if constexpr (FIN == FIN_CG) load_partials(pvr, a.part_in + pb0, pb1 - pb0);

  // first PF phases in flight before the vector staging
  PhaseSlot<T, NDP, ncoef> pipe[PF > 0 ? PF : 1];
  if constexpr (PF > 0) {
#pragma unroll
    for (int q = 0; q < PF; ++q)
      load_phase<T, TA, NDP, ncoef>(pipe[q], q, sh_co, se, a.lid, a.xa, a.nslot_pad,
                                    a.col_ids);
  }

  const int n0 = a.node_start[b], n1 = a.node_start[b + 1];
  const int nown = n1 - n0;
  const int h0 = a.halo_start[b], h1 = a.halo_start[b + 1];
  const int nloc = nown + (h1 - h0);

  // ---- early exits of the CG kernels (state written by earlier kernels) --
  if constexpr (FIN == FIN_CG) {
    SolveState* st = a.st + sys;
    if (a.skip_unless_fresh && !st->fresh) return;
    if (load_relaxed(&st->stop)) return;
    if (load_relaxed(&st->final_flag)) {
      // the update kernel before this one applied the last X update
      if (b == pb0 && threadIdx.x == 0) {
        st->t_done = wall_clock64();
        store_relaxed(&st->done, ST_MAXITER);
        store_relaxed(&st->done1, (st->final_m << 2) | ST_MAXITER);
      }
      return;
    }
  }

  // ---- node values of the staging, loaded before the scalar reduction ----
  // (the first kStagePre nodes of each thread) so that their HBM latency
  // overlaps the reduction of the previous kernel's partial sums
  constexpr bool kPre = STAGE == STAGE_CGDIR;
  constexpr int KS = kPre ? (sizeof(T) == 4 ? kStagePreF32 : kStagePre) : 1;
  NodeIn<T> pre[KS];
  auto load_node = [&](int l, NodeIn<T>& v) {
    const bool owned = l < nown;
    v.g = owned ? n0 + l : a.halo[h0 + l - nown];
    if constexpr (STAGE == STAGE_CGDIR) {
      v.r = a.r[v.g];
      v.d = a.dold[v.g];
    }
  };
  if constexpr (kPre) {
#pragma unroll
    for (int k = 0; k < KS; ++k) {
      const int l = threadIdx.x + k * kThreads;
      if (l < nloc) load_node(l, pre[k]);
    }
  }

  double cg_beta = 0.0;
  if constexpr (FIN == FIN_CG) {
    // E_m: RMRM from the update kernel's partials, stop test (gracjg.F:306-324,
    // and the strict test of the initial residual, gracjg.F:219), BETA
    SolveState* st = a.st + sys;
    const int p = a.parity;
    const double rr = sum_partials(pvr, a.part_in + pb0, pb1 - pb0, sh_red, &sh_b);
    const int m = st->m[1 - p] + 1;
    const double thr = st->use_rr0 ? st->eps2xl * st->rr0 : st->eps2xl;
    const bool conv = (m == 0) ? (rr < thr) : (rr <= thr);
    if (conv) {
      if (b == pb0 && threadIdx.x == 0) {
        st->final_m = m;
        st->final_rr = rr;
        st->t_done = wall_clock64();
        store_relaxed(&st->done, ST_CONVERGED);
        store_relaxed(&st->done1, (m << 2) | ST_CONVERGED);
      }
      return;
    }
    cg_beta = (m == 0) ? 0.0 : rr / st->rr[1 - p];
    if (b == pb0 && threadIdx.x == 0) {
      st->rr[p] = rr;
      st->m[p] = m;
      st->status[p] = (m >= st->nitmax) ? ST_FINAL : ST_RUNNING;
    }
  }

  T* ys = reinterpret_cast<T*>(smem_raw);
  T* acc = ys + ((nloc + 1) & ~1);

  // ---- stage the multiplied vector (owned + halo) into LDS ---------------
  auto put = [&](int l, int g, T v) {
    ys[l] = v;
    // diagonal term (mv0303.f:150-160)
    if (l < nown) {
      T dv;
      if constexpr (DIAG == DIAG_Q) dv = (T)a.da[g] * v;
      else if constexpr (DIAG == DIAG_ONE) dv = v;
      else dv = (T)0;
      acc[l] = dv;
    }
  };
  if constexpr (kPre) {
    auto stage_node = [&](int l, const NodeIn<T>& v) {
      const bool owned = l < nown;
      const int g = v.g;
      const T val = v.r + (T)cg_beta * v.d;
      if (owned) {
        a.dnew[g] = val;
        if (a.xzero) a.xzero[g] = (T)0;
      }
      put(l, g, val);
    };
#pragma unroll
    for (int k = 0; k < KS; ++k) {
      const int l = threadIdx.x + k * kThreads;
      if (l < nloc) stage_node(l, pre[k]);
    }
    for (int l0 = threadIdx.x + KS * kThreads; l0 < nloc; l0 += kStageUnroll * kThreads) {
      NodeIn<T> v[kStageUnroll];
#pragma unroll
      for (int u = 0; u < kStageUnroll; ++u) {
        const int l = l0 + u * kThreads;
        if (l < nloc) load_node(l, v[u]);
      }
#pragma unroll
      for (int u = 0; u < kStageUnroll; ++u) {
        const int l = l0 + u * kThreads;
        if (l < nloc) stage_node(l, v[u]);
      }
    }
  } else {
    // STAGE_XUPD: x - delta from the system's current FP64 buffer, written
    // to the other one (halo nodes of other blocks read the current buffer,
    // which no workgroup of this kernel writes)
    const double* xcur = nullptr;
    double* xnxt = nullptr;
    if constexpr (STAGE == STAGE_XUPD) {
      const int xp = a.st[sys].xpar;
      xcur = xp ? a.xb1 : a.xb0;
      xnxt = xp ? a.xb0 : a.xb1;
    }
    for (int l = threadIdx.x; l < nloc; l += blockDim.x) {
      const bool owned = l < nown;
      const int g = owned ? n0 + l : a.halo[h0 + l - nown];
      T v;
      if constexpr (STAGE == STAGE_PLAIN) {
        v = a.y[g];
      } else {  // STAGE_XUPD
        const double xv = xcur[g] - (double)a.delta[g];
        v = (T)xv;
        if (owned) xnxt[g] = xv;
      }
      put(l, g, v);
    }
  }
  __syncthreads();
  // ---- elements, colour by colour ----------------------------------------
  if constexpr (PF > 0) {
    for (int c0 = 0; c0 < kMaxColors; c0 += PF) {
      if (sh_co[c0] >= se) break;
#pragma unroll
      for (int q = 0; q < PF; ++q) {
        const int c = c0 + q;
        if (c < kMaxColors && sh_co[c] < se) {
          if (pipe[q].valid) slot_apply<T, NDP, SYMQ, TRANS>(pipe[q].l, pipe[q].cf, ys, acc, nown);
          // slots beyond the first kThreads of a large phase
          for (int s = sh_co[c] + (int)threadIdx.x + kThreads; s < sh_co[c + 1];
               s += kThreads) {
            int l[NDP];
            T cf[ncoef];
            load_ids<NDP>(l, a.lid, a.nslot_pad, s, a.col_ids);
#pragma unroll
            for (int qq = 0; qq < ncoef; ++qq) cf[qq] = (T)a.xa[(size_t)qq * a.nslot_pad + s];
            slot_apply<T, NDP, SYMQ, TRANS>(l, cf, ys, acc, nown);
          }
          // refill this pipeline stage with phase c + PF before the barrier
          load_phase<T, TA, NDP, ncoef>(pipe[q], c + PF, sh_co, se, a.lid, a.xa, a.nslot_pad,
                                        a.col_ids);
        }
        __syncthreads();
      }
    }
  } else {
    for (int c = 0; c < kMaxColors; ++c) {
      const int c0 = sh_co[c], c1 = sh_co[c + 1];
      if (c0 >= se) break;
      for (int s = c0 + threadIdx.x; s < c1; s += blockDim.x) {
        int l[NDP];
        T cf[ncoef];
        load_ids<NDP>(l, a.lid, a.nslot_pad, s, a.col_ids);
#pragma unroll
        for (int q = 0; q < ncoef; ++q) cf[q] = (T)a.xa[(size_t)q * a.nslot_pad + s];
        slot_apply<T, NDP, SYMQ, TRANS>(l, cf, ys, acc, nown);
      }
      __syncthreads();
    }
  }
  // ---- epilogue at owned nodes -------------------------------------------
  double part = 0.0, part2 = 0.0;
  for (int l = threadIdx.x; l < nown; l += blockDim.x) {
    const int g = n0 + l;
    const T v = acc[l];
    if constexpr (EPI == EPI_OP) {
      T o;
      switch (a.op) {
        case BHIP_OP_AY: case BHIP_OP_TAY: o = v; break;
        case BHIP_OP_CAY: o = (T)a.c * v; break;
        case BHIP_OP_NAY: case BHIP_OP_NTAY: o = -v; break;
        case BHIP_OP_XPAY: case BHIP_OP_XPTAY: o = a.xin[g] + v; break;
        case BHIP_OP_XMAY: case BHIP_OP_XMTAY: o = a.xin[g] - v; break;
        default: o = a.xin[g] + (T)a.c * v; break;  // X+CAY, X+CTAY
      }
      a.out[g] = o;
    } else if constexpr (EPI == EPI_RESID) {
      const double rv = (double)v - a.b[g];
      a.out[g] = (T)rv;
      part += rv * rv;
    } else if constexpr (EPI == EPI_CGAD) {
      a.out[g] = v;
      part += (double)ys[l] * (double)v;
    } else if constexpr (EPI == EPI_DOTW) {
      a.out[g] = v;
      part += (double)a.w[g] * (double)v;
    } else if constexpr (EPI == EPI_DOT2) {
      a.out[g] = v;
      part += (double)v * (double)ys[l];
      part2 += (double)v * (double)v;
    } else {  // EPI_RESID_MIXED: FP64 residual, inner RHS in FP32
      const double rv = (double)v - a.b[g];
      a.r32[g] = (float)(-rv);
      part += rv * rv;
    }
  }
  if constexpr (FIN != FIN_NONE) {
    double s = block_sum(part, sh_red);
    if (threadIdx.x == 0) a.partial[b] = s;
    if constexpr (EPI == EPI_DOT2) {
      double s2 = block_sum(part2, sh_red);
      if (threadIdx.x == 0) a.partial2[b] = s2;
    }
  }
}

// ---------------------------------------------------------------------------
// CG vector update: X -= RO D ; R -= RO AD ; RMRM = R.R (gracjg.F:284-296,
// 387-388), with the state transitions of the loop test (gracjg.F:306-324)
// ---------------------------------------------------------------------------
constexpr int kUpdPre = 4;   // nodes per thread loaded ahead in k_cg_update
constexpr int kUpdPart = 8;  // partial sums per thread loaded ahead

template <typename T, typename TX>
__global__ void __launch_bounds__(kThreads)
k_cg_update(TX* __restrict__ x, T* __restrict__ r, const T* __restrict__ d,
            const T* __restrict__ ad, const int32_t* __restrict__ node_start, int nblocks,
            const int32_t* __restrict__ block_sys, const int32_t* __restrict__ sys_blocks,
            const double* __restrict__ part_e, double* __restrict__ part_u, SolveState* st0,
            int parity) {
  __shared__ double sh_red[kWaves];
  __shared__ double sh_b;
  // one workgroup per node block, in the direction kernel's block order (its
  // AD and D of these nodes were written through the same XCD's L2)
  const int b = xcd_block(blockIdx.x, nblocks);
  const int sys = block_sys ? block_sys[b] : 0;
  const int pb0 = sys_blocks ? sys_blocks[sys] : 0;
  const int pb1 = sys_blocks ? sys_blocks[sys + 1] : nblocks;
  SolveState* st = st0 + sys;
  if (load_relaxed(&st->done)) {
    // the direction kernel before this one found the stop: later kernels skip
    if (b == pb0 && threadIdx.x == 0) store_relaxed(&st->stop, 1);
    return;
  }
  // the first update of an inner solve retires its start (no workgroup of
  // this kernel reads the flag)
  if (b == pb0 && threadIdx.x == 0 && st->fresh) st->fresh = 0;
  const int n0 = node_start[b], n1 = node_start[b + 1];
  const int q = 1 - parity;  // slot of iteration m - 1
  const int status = st->status[q];
  const double rrq = st->rr[q];
  // loads first, arithmetic after: this thread's share of the partial sums
  // of D.AD, then its first kUpdPre nodes, so that the HBM latency of the
  // node data overlaps the reduction (the sum keeps reduce_partials' order).
  // Indices are clamped rather than branched on, so no load waits on another.
  const double* pe = part_e + pb0;
  const int np = pb1 - pb0;
  double pv[kUpdPart];
  load_partials(pv, pe, np);
  // (FP64 keeps fewer nodes in flight: its registers would cost occupancy)
  constexpr int kPre = sizeof(T) == 4 ? kUpdPre : kUpdPre / 2;
  T dv[kPre], adv[kPre], rv0[kPre];
  TX xv[kPre];
#pragma unroll
  for (int k = 0; k < kPre; ++k) {
    const int i = min(n0 + (int)threadIdx.x + k * kThreads, n1 - 1);
    dv[k] = d[i];
    adv[k] = ad[i];
    rv0[k] = r[i];
    xv[k] = x[i];
  }
  // RO = RMRM / D.AD of iteration m - 1 (gracjg.F:380-383)
  const double dad = sum_partials(pv, pe, np, sh_red, &sh_b);
  const double rho = rrq / dad;
  const T rhot = (T)rho;
  if (status == ST_FINAL) {
    // NITMAX reached: the last X = X - RO D, then stop (gracjg.F:387-409)
    for (int i = n0 + threadIdx.x; i < n1; i += blockDim.x) x[i] = x[i] - (TX)rhot * (TX)d[i];
    // every workgroup of the system runs this branch (status[q] was written
    // by an earlier kernel); the next direction kernel turns final_flag into done
    if (b == pb0 && threadIdx.x == 0) {
      st->final_m = st->m[q];
      st->final_rr = st->rr[q];
      st->rho[q] = rho;
      store_relaxed(&st->final_flag, 1);
    }
    return;
  }
  double part = 0.0;
#pragma unroll
  for (int k = 0; k < kPre; ++k) {
    const int i = n0 + threadIdx.x + k * kThreads;
    if (i < n1) {
      x[i] = xv[k] - (TX)rhot * (TX)dv[k];
      const T rv = rv0[k] - rhot * adv[k];
      r[i] = rv;
      part += (double)rv * (double)rv;
    }
  }
  for (int i = n0 + threadIdx.x + kPre * kThreads; i < n1; i += blockDim.x) {
    x[i] = x[i] - (TX)rhot * (TX)d[i];
    const T rv = r[i] - rhot * ad[i];
    r[i] = rv;
    part += (double)rv * (double)rv;
  }
  double sm = block_sum(part, sh_red);
  if (threadIdx.x == 0) part_u[b] = sm;
  if (b == pb0 && threadIdx.x == 0) st->rho[q] = rho;
}

// ---------------------------------------------------------------------------
// small kernels
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_dot_partials(const T* x, const T* y, int n, double* partial) {
  __shared__ double sh_red[kWaves];
  double part = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    part += (double)x[i] * (double)y[i];
  double s = block_sum(part, sh_red);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// partial[b] = x.y over the nodes of block b (one workgroup per block)
__global__ void k_dot_blocks(const double* __restrict__ x, const double* __restrict__ y,
                             const int32_t* __restrict__ node_start, double* partial) {
  __shared__ double sh_red[kWaves];
  const int b = blockIdx.x;
  double part = 0.0;
  for (int i = node_start[b] + threadIdx.x; i < node_start[b + 1]; i += blockDim.x)
    part += x[i] * y[i];
  double s = block_sum(part, sh_red);
  if (threadIdx.x == 0) partial[b] = s;
}

// one workgroup: *dst = fixed-order sum of n partials
__global__ void k_reduce_into(const double* __restrict__ partial, int n, double* dst) {
  __shared__ double sh_red[kWaves];
  __shared__ double sh_b;
  double s = reduce_partials(partial, n, sh_red, &sh_b);
  if (threadIdx.x == 0) *dst = s;
}

// an inner solve has applied its last update
__device__ inline bool inner_finished(const SolveState* st) {
  return load_relaxed(&st->done1) != 0;
}

// the mixed outer step of a system waits: stopped, or its inner solve still running
__device__ inline bool outer_blocked(const SolveState* st) {
  return load_relaxed(&st->outer_done) ||
         (load_relaxed(&st->outer_count) > 0 && !inner_finished(st));
}

__device__ inline void reset_cg(SolveState* st, double eps2xl, int use_rr0, int nitmax) {
  st->eps2xl = eps2xl;
  st->use_rr0 = use_rr0;
  st->nitmax = nitmax < 1 ? 1 : nitmax;
  st->m[0] = st->m[1] = -1;
  st->status[0] = st->status[1] = ST_RUNNING;
  st->rr[0] = st->rr[1] = 1.0;
  st->final_flag = 0;
  st->final_m = 0;
  st->final_rr = 0.0;
  st->done = 0;
  st->done1 = 0;
  st->stop = 0;
}

// GRACJG set-up (gracjg.F:189-209) from the partial sums of ||b'||^2: XL,
// the zero right-hand-side exit, the relative/absolute switch; then the
// state of a fresh CG solve (FP64) or of the mixed outer loop.
// One workgroup per system; partial holds one sum per node block.
__global__ void k_cg_setup(const double* __restrict__ partial, int nblocks,
                           const int32_t* __restrict__ sys_blocks, double eps, int nitmax,
                           int mixed, double inner_eps, int inner_nitmax, int reliable,
                           SolveState* st0) {
  __shared__ double sh_red[kWaves];
  __shared__ double sh_b;
  const int sys = blockIdx.x;
  SolveState* st = st0 + sys;
  const int pb0 = sys_blocks ? sys_blocks[sys] : 0;
  const int pb1 = sys_blocks ? sys_blocks[sys + 1] : nblocks;
  const double xl_raw = reduce_partials(partial + pb0, pb1 - pb0, sh_red, &sh_b);
  if (threadIdx.x != 0) return;
  const double xl = xl_raw < 1.0 ? 1.0 : xl_raw;
  st->xl_raw = xl_raw;
  st->xl = xl;
  st->zero_rhs = sqrt(xl_raw) < 1e-15;
  st->outer_done = 0;
  st->outer_status = 0;
  st->outer_count = 0;
  st->inner_total = 0;
  st->rr_outer = 0.0;
  st->dpar = -1;
  st->fresh = 0;
  st->xpar = 0;
  st->reliable = mixed && reliable;
  for (int k = 0; k < 8; ++k) st->inner_pairs[k] = 0;
  st->t_start = wall_clock64();
  st->t_done = 0;
  if (!mixed) {
    reset_cg(st, eps * eps * xl, 0, nitmax);
  } else {
    st->eps2_outer = eps * eps * xl;
    st->inner_eps2 = inner_eps * inner_eps;
    st->inner_nitmax = inner_nitmax < 1 ? 1 : inner_nitmax;
    st->nitmax_total = nitmax < 1 ? 1 : nitmax;
    reset_cg(st, inner_eps * inner_eps, 1, inner_nitmax);
    st->done = ST_CONVERGED;  // no inner solve until an outer test starts one
    st->stop = 1;
  }
  if (st->zero_rhs) {
    st->done = ST_CONVERGED;
    st->stop = 1;
    st->outer_done = 1;
    st->outer_status = ST_CONVERGED;
    st->final_m = 0;
    st->final_rr = 0.0;
    st->t_done = st->t_start;
  }
}

// Mixed precision outer test, one workgroup per system: FP64 ||r'||^2 from
// the partials of the system's blocks in the residual kernel; converged or
// out of iterations -> the system stops, else the next inner solve starts
// (its state reset on the device, rr0 = ||r'||^2). A system whose inner
// solve is still running had no residual this step and is left alone, as
// the residual kernel left it (outer_blocked).
constexpr int kMaxOuter = 100;  // refinement sweeps of one mixed solve
__global__ void k_outer_check(const double* __restrict__ partial, int nblocks,
                              const int32_t* __restrict__ sys_blocks, SolveState* st0,
                              int flip) {
  __shared__ double sh_red[kWaves];
  __shared__ double sh_b;
  const int sys = blockIdx.x;
  SolveState* st = st0 + sys;
  if (outer_blocked(st)) return;
  const int pb0 = sys_blocks ? sys_blocks[sys] : 0;
  const int pb1 = sys_blocks ? sys_blocks[sys + 1] : nblocks;
  const double rr = reduce_partials(partial + pb0, pb1 - pb0, sh_red, &sh_b);
  if (threadIdx.x != 0) return;
  // every field this kernel reads, read once before any store, with
  // agent-scope loads: a plain load of a field written by an earlier kernel
  // has returned a stale value here (depending on how the compiler issued it)
  const int oc = load_relaxed(&st->outer_count);
  const int fm = load_relaxed(&st->final_m);
  const int dp = load_relaxed(&st->dpar);
  const int dn = load_relaxed(&st->done);
  const int rel = load_relaxed(&st->reliable);
  const int itot0 = load_relaxed(&st->inner_total);
  const int ntot = load_relaxed(&st->nitmax_total);
  const int innmax = load_relaxed(&st->inner_nitmax);
  const int xp = load_relaxed(&st->xpar);
  const double eps2o = st->eps2_outer;
  const double ie2 = st->inner_eps2;
  const double rr_s0 = st->rr[0], rr_s1 = st->rr[1];
  if (flip) store_relaxed(&st->xpar, xp ^ 1);
  int itot = itot0;
  if (oc > 0) {
    const int used = fm - (dp >= 0 ? 1 : 0);
    itot += used;
    if (oc <= 8) st->inner_pairs[oc - 1] = used;
  }
  st->inner_total = itot;
  st->rr_outer = rr;
  const int left = ntot - itot;
  if (rr <= eps2o || left <= 0 || oc >= kMaxOuter) {
    st->outer_status = rr <= eps2o ? ST_CONVERGED : ST_MAXITER;
    st->outer_done = 1;
    st->done = ST_CONVERGED;
    st->stop = 1;
    st->t_done = wall_clock64();
    return;
  }
  const int m0 = dp >= 0 ? 1 : 0;
  const int slot = (fm - 1 - m0) & 1;
  const bool cont = rel && oc > 0 && fm - 1 >= m0 && dn == ST_CONVERGED;
  const double rprev = cont ? (slot ? rr_s1 : rr_s0) : 1.0;
  st->outer_count = oc + 1;
  const int cap = (innmax < left ? innmax : left) + (cont ? 1 : 0);
  reset_cg(st, ie2, 1, cap);
  st->rr0 = rr;
  st->dpar = -1;
  if (cont) {
    st->m[1] = 0;
    st->rr[1] = rprev;
    st->dpar = slot;
  }
  st->fresh = 1;
}

// the direction a continued inner solve starts from, copied to dst (the
// start kernel reads it while writing d0, so it cannot stay in d0); one
// workgroup per node block, for the systems whose inner solve starts
__global__ void k_inner_dir(float* __restrict__ dst, const float* __restrict__ d0,
                            const float* __restrict__ d1, const int32_t* __restrict__ node_start,
                            const int32_t* __restrict__ block_sys, const SolveState* st0) {
  const int b = blockIdx.x;
  const SolveState* st = st0 + (block_sys ? block_sys[b] : 0);
  const int par = load_relaxed(&st->dpar);
  if (!load_relaxed(&st->fresh) || par < 0 || load_relaxed(&st->outer_done)) return;
  const float* src = par ? d1 : d0;
  for (int i = node_start[b] + threadIdx.x; i < node_start[b + 1]; i += blockDim.x)
    dst[i] = src[i];
}

// the solution into buffer 0 for the systems whose last outer step left it
// in buffer 1 (one workgroup per node block)
__global__ void k_xpar_gather(double* __restrict__ x0, const double* __restrict__ x1,
                              const int32_t* __restrict__ node_start,
                              const int32_t* __restrict__ block_sys, const SolveState* st0) {
  const int b = blockIdx.x;
  const SolveState* st = st0 + (block_sys ? block_sys[b] : 0);
  if (!load_relaxed(&st->xpar)) return;
  for (int i = node_start[b] + threadIdx.x; i < node_start[b + 1]; i += blockDim.x)
    x0[i] = x1[i];
}

// out[i] = in[idx[i]]
template <typename T>
__global__ void k_gather(T* __restrict__ out, const T* __restrict__ in,
                         const int32_t* __restrict__ idx, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    out[i] = in[idx[i]];
}
// out[idx[i]] = in[i]
template <typename T>
__global__ void k_scatter(T* __restrict__ out, const T* __restrict__ in,
                          const int32_t* __restrict__ idx, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    out[idx[i]] = in[i];
}
// SOLVE's results back to the caller: out[idx[i]] (out[i] without idx) =
// src[i] (* s[i]), src the solution buffer the system's state names
// (st: in1 when xpar is set, else in0; no state: in0)
__global__ void k_x_out(double* __restrict__ out, const double* __restrict__ in0,
                        const double* __restrict__ in1, const SolveState* st,
                        const double* __restrict__ s, const int32_t* __restrict__ idx, int n) {
  const double* src = (st && load_relaxed(&st->xpar)) ? in1 : in0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double v = s ? src[i] * s[i] : src[i];
    out[idx ? idx[i] : i] = v;
  }
}
template <typename T>
__global__ void k_fill(T* __restrict__ x, T v, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    x[i] = v;
}
__global__ void k_to_f32(float* __restrict__ out, const double* __restrict__ in, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    out[i] = (float)in[i];
}

// PRECD1 node part (precd1.f:90-149): D = sqrt(DA) (|DA| for PRECON 5);
// X = X*D; D = 1/D guarded (|D| > 1e-10, else 1: ovd.f:92-102); B = B*D;
// DA = DA*D*D ('Q') or D*D ('I') (om1111.f:559-575).
// X and B are read from the caller's vectors (xs, bs; through perm when it
// is given, i.e. in the caller's node order) and written to the context's
// internal-order copies x and b, so SOLVE needs no separate gather.
// da_in and da_out may be the same array (the context's diagonal in place),
// or the saved copy and the context's diagonal (OM 'M=N' fused into SOLVE)
__global__ void k_precd1_nodes(double* x, double* b, const double* xs, const double* bs,
                               const int32_t* __restrict__ perm, const double* da_in,
                               double* da_out, double* __restrict__ dout, int n, int precon,
                               int typdia_q, int typdia_i) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int j = perm ? perm[i] : i;
    const double xv = xs[j], bv = bs[j];
    const double a = da_in[i];
    double d = a;
    if (precon == 5) d = fabs(d);
    d = sqrt(d);
    x[i] = xv * d;
    d = (fabs(d) > 1e-10) ? 1.0 / d : 1.0;
    dout[i] = d;
    if (typdia_q) {
      const double v = a * d;
      da_out[i] = v * d;
    } else if (typdia_i) {
      da_out[i] = d * d;
    } else if (da_out != da_in) {
      da_out[i] = a;
    }
    b[i] = bv * d;
  }
}

// OM 'M=DMD' on element coefficients: XM(e,c) = XM(e,c)*D(IKLE(e,j))*D(IKLE(e,i))
// for the pair (i<j) of column c (om1111.f:583-600, om4141.f), per slot. One
// workgroup per node block: D of its owned and halo nodes is staged in LDS,
// so the per-slot gathers are LDS reads. xa32 (optional) receives an FP32
// copy of the result. The coefficients are read from src, which is xa itself
// or the saved copy (OM 'M=N' fused into SOLVE); each slot's are all loaded
// before any is stored.
template <int NDP>
__global__ void __launch_bounds__(kThreads)
k_dmd_slots(const double* src, double* xa, float* __restrict__ xa32, int ncoef, int nslot_pad,
            const uint16_t* __restrict__ lid, const double* __restrict__ d,
            const int32_t* __restrict__ node_start, const int32_t* __restrict__ halo_start,
            const int32_t* __restrict__ halo, const int32_t* __restrict__ color_off,
            int nblocks) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double* ds = reinterpret_cast<double*>(smem_raw);
  const int nb = nblocks;
  int b;
  {
    const int g = blockIdx.x;
    const int q = nb >> 3, r = nb & 7;
    const int xcd = g & 7, k = g >> 3;
    b = xcd * q + (xcd < r ? xcd : r) + k;
  }
  const int n0 = node_start[b], nown = node_start[b + 1] - n0;
  const int h0 = halo_start[b], nloc = nown + (halo_start[b + 1] - h0);
  for (int l = threadIdx.x; l < nloc; l += blockDim.x)
    ds[l] = d[l < nown ? n0 + l : halo[h0 + l - nown]];
  __syncthreads();
  const int s0 = color_off[(size_t)b * (kMaxColors + 1)];
  const int s1 = color_off[(size_t)b * (kMaxColors + 1) + kMaxColors];
  constexpr int npair = NDP * (NDP - 1) / 2;
  const bool two = ncoef > npair;
  for (int s = s0 + threadIdx.x; s < s1; s += blockDim.x) {
    double dv[NDP], c1[npair], c2[npair];
#pragma unroll
    for (int k = 0; k < NDP; ++k) dv[k] = ds[lid[(size_t)k * nslot_pad + s]];
#pragma unroll
    for (int q = 0; q < npair; ++q) {
      c1[q] = src[(size_t)q * nslot_pad + s];
      c2[q] = two ? src[(size_t)(q + npair) * nslot_pad + s] : 0.0;
    }
    int q = 0;
#pragma unroll
    for (int i = 0; i < NDP; ++i)
#pragma unroll
      for (int j = i + 1; j < NDP; ++j) {
        // elements: XM * D(j) * D(i) (om1111.f:583-600); segments (NDP 2):
        // XM * D(GLOSEG1) * D(GLOSEG2) (omseg.f:524-536)
        const double v1 = NDP == 2 ? c1[q] * dv[i] * dv[j] : c1[q] * dv[j] * dv[i];
        xa[(size_t)q * nslot_pad + s] = v1;
        if (xa32) xa32[(size_t)q * nslot_pad + s] = (float)v1;
        if (two) {
          const double v2 = NDP == 2 ? c2[q] * dv[i] * dv[j] : c2[q] * dv[j] * dv[i];
          xa[(size_t)(q + npair) * nslot_pad + s] = v2;
          if (xa32) xa32[(size_t)(q + npair) * nslot_pad + s] = (float)v2;
        }
        ++q;
      }
  }
}

// dst[i] = src[i], i < n (16-byte vectors; both 16-byte aligned)
__global__ void k_copy(double* __restrict__ dst, const double* __restrict__ src, long n) {
  const long n2 = n >> 1;
  double2* d2 = reinterpret_cast<double2*>(dst);
  const double2* s2 = reinterpret_cast<const double2*>(src);
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n2;
       i += (long)gridDim.x * blockDim.x)
    d2[i] = s2[i];
  if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) dst[n - 1] = src[n - 1];
}

// xa_slots[c][s] = xa_caller[c][slot_elem[s]]
__global__ void k_matrix_to_slots(double* __restrict__ dst, const double* __restrict__ src,
                                  const int32_t* __restrict__ slot_elem, int nslot,
                                  int nslot_pad, int nelem, int ncoef) {
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < (long)nslot * ncoef;
       t += (long)gridDim.x * blockDim.x) {
    const int c = (int)(t / nslot), s = (int)(t % nslot);
    dst[(size_t)c * nslot_pad + s] = src[(size_t)c * nelem + slot_elem[s]];
  }
}
// caller layout back from slots: first instance of each element wins (all
// instances of an element hold identical values)
__global__ void k_matrix_from_slots(double* __restrict__ dst, const double* __restrict__ src,
                                    const int32_t* __restrict__ first_slot, int nelem,
                                    int nslot_pad, int ncoef) {
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < (long)nelem * ncoef;
       t += (long)gridDim.x * blockDim.x) {
    const int c = (int)(t / nelem), e = (int)(t % nelem);
    dst[(size_t)c * nelem + e] = src[(size_t)c * nslot_pad + first_slot[e]];
  }
}

// ---------------------------------------------------------------------------
// host side: context
// ---------------------------------------------------------------------------
// Device buffer owned by its holder: freed by release() or when it goes out
// of scope (temporaries of an entry point, whose work has been waited for).
template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  hipError_t alloc(size_t count) {
    if (p && n >= count) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    hipError_t e = hipMalloc((void**)&p, std::max<size_t>(count, 1) * sizeof(T));
    if (e == hipSuccess) n = count;
    return e;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
};

}  // namespace bhip

struct bhip_ctx {
  bhip::Layout L;
  hipStream_t stream = nullptr;
  int device = 0;
  int grid_nodes = 0;  // grid of the node-wise kernels
  // layout on device
  bhip::DevBuf<int32_t> node_start, halo_start, halo, color_off, slot_elem, perm, first_slot;
  bhip::DevBuf<uint16_t> lid;
  // matrix
  char typdia = '0', typext = '0';
  int ncoef = 0;
  bool have_matrix = false;
  bhip::DevBuf<double> da, xa;       // internal node order / slot order
  bhip::DevBuf<float> da32, xa32;
  // work vectors (internal order)
  bhip::DevBuf<double> vx, vb, vr, vd0, vd1, vad, vdz, vdinv, vio0, vio1;
  bhip::DevBuf<double> vx2;      // mixed path: second FP64 solution buffer
  bhip::DevBuf<float> fr, fd0, fd1, fad, fx0, fdz;
  bhip::DevBuf<double> partial;  // 4 arrays of max_part: U[0], U[1], E[0], E[1]
  int max_part = 0;
  bhip::DevBuf<double> geo;      // element geometry per slot (bhip_ctx_set_geometry)
  bool have_geo = false;
  // inputs of the operand-preparation entry points in the context's order:
  // node fields (7 slots of npoin: six inputs and a host staging area) and
  // element fields (4 slots of nelem); in_dev: the current call's arrays are
  // device pointers (the _dev entry points), so nothing is copied from the
  // host and nothing waits
  bhip::DevBuf<double> nsc, esc;
  bhip::DevBuf<int32_t> isc;
  bool in_dev = false;
  bool reliable = true;          // mixed inner CG continued across outer steps
  bhip::DevBuf<bhip::SolveState> st;  // one per system
  // the partial sums and solver states the kernels use: the context's own
  // buffers above, or the library's static scratch during a call of
  // bhip_solve_dev_noalloc
  double* part_p = nullptr;
  bhip::SolveState* st_p = nullptr;
  // ensembles of independent meshes (nsys > 1): the system of each node
  // block, the block range and node/element offsets of each system
  int nsys = 1;
  bhip::DevBuf<int32_t> block_sys, sys_blocks;
  std::vector<int32_t> sys_blocks_h;
  std::vector<int64_t> sys_node_off, sys_elem_off;
  std::vector<bhip_solve_info> sys_info;  // per system, last solve
  std::vector<double> sys_latency_ms;     // per system, last CG solve: device time
                                          // from its set-up to its stop
  // captured CG iteration batches (key: precision, solution buffer)
  struct Graph {
    const void* key = nullptr;
    int prec = -1;
    int tag = 0;  // TYPDIA, TYPEXT the kernels were specialised for
    hipGraphExec_t exec = nullptr;
  };
  Graph graphs[4];
  // the planned part of whole solves (set-up, expected iteration batches),
  // captured once per plan and replayed as one graph
  struct SolveGraph {
    uint64_t key[8];
    hipGraphExec_t exec;
  };
  std::vector<SolveGraph> solve_graphs;
  bool use_graphs = true;
  int prefetch_j = 0;  // register prefetch depth of the element kernels
  int iter_hint[4] = {0, 0, 0, 0};  // FP64 CG, inner CG, inner solves, CGSTAB
  int inner_hint[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // iterations of inner solve k (mixed CG)
  // BiCGSTAB scalars and extra partial sums; GMRES basis and scalars
  bhip::DevBuf<double> bicg_sc, bicg_part, gm, gm_scal;
  bhip::DevBuf<float> gm32;      // mixed GMRES: FP32 basis V and AV
  // a saved copy of the matrix (OM 'M=N' target), slot layout
  bhip::DevBuf<double> da_saved, xa_saved;
  char typdia_saved = '0', typext_saved = '0';
  int ncoef_saved = 0;
  bool have_saved = false;
  bhip::SolveState* st_host = nullptr;  // pinned
  size_t lds_f64 = 0, lds_f32 = 0;
  // the time-stepped wave equation on the device (wave.inl): slot geometry
  // [5][nslot_pad] (x2, x3, y2, y3, S of the slot's element) and node fields
  // in internal order (1 - sea mask, basis integrals, h, c, qx, qy, y, and
  // the assembled diagonal before preconditioning)
  bhip::DevBuf<double> wv_geo, wv_w, wv_vol, wv_h, wv_zb, wv_f, wv_vs, wv_c, wv_usx, wv_usy,
      wv_fb, wv_da, wv_sbx, wv_sby, wv_svs, wv_seb;
  bool wv_ready = false;
  // device copies of a host entry point's caller arrays (kept between calls)
  bhip::DevBuf<double> h_xa, h_x, h_b;
};

#include "hostcopy.inl"

namespace bhip {

// the per-system arrays of an ensemble; NULL for one system
static const int32_t* sys_ptr(const DevBuf<int32_t>& v) { return v.n ? v.p : nullptr; }

static size_t lds_bytes(const Layout& L, size_t elt) {
  size_t nl = ((size_t)L.max_local + 1) & ~(size_t)1;
  return (nl + (size_t)L.max_owned) * elt;
}

// LDS of the DMD kernel: D of a block's owned and halo nodes
static size_t dmd_lds(const Layout& L) { return ((size_t)L.max_local + 1) * sizeof(double); }

static int grid_for(int n);

// OM 'M=DMD' on the context's coefficients with the node vector d (and the
// FP32 copy of the result when xa32 is given); the coefficients come from src
// (default: the context's own)
static hipError_t launch_dmd(bhip_ctx* c, float* xa32, const double* d,
                             const double* src = nullptr) {
  const Layout& L = c->L;
  if (!src) src = c->xa.p;
#define BHIP_DMD(NDP)                                                                          \
  hipLaunchKernelGGL(k_dmd_slots<NDP>, dim3(L.nblocks), dim3(kThreads), dmd_lds(L), c->stream, \
                     src, c->xa.p, xa32, c->ncoef, L.nslot_pad, c->lid.p, d, c->node_start.p, \
                     c->halo_start.p, c->halo.p, c->color_off.p, L.nblocks)
  if (L.ndp == 2)
    BHIP_DMD(2);
  else if (L.ndp == 3)
    BHIP_DMD(3);
  else
    BHIP_DMD(6);
#undef BHIP_DMD
  return hipGetLastError();
}

static int grid_for(int n) {
  int g = (n + kThreads * 8 - 1) / (kThreads * 8);
  return std::max(1, std::min(g, 1024));
}

template <typename T, typename TA>
static EbeArgs<T, TA> base_args(bhip_ctx* c, const TA* xa, const TA* da) {
  EbeArgs<T, TA> a;
  std::memset(&a, 0, sizeof(a));
  a.node_start = c->node_start.p;
  a.halo_start = c->halo_start.p;
  a.halo = c->halo.p;
  a.color_off = c->color_off.p;
  a.lid = c->lid.p;
  a.col_ids = c->L.col_ids ? 1 : 0;
  a.nslot_pad = c->L.nslot_pad;
  a.nblocks = c->L.nblocks;
  a.xa = xa;
  a.da = da;
  a.partial = c->part_p;
  a.st = c->st_p;
  a.block_sys = sys_ptr(c->block_sys);
  a.sys_blocks = sys_ptr(c->sys_blocks);
  return a;
}

// dispatch over (ndp, symmetry, transpose, diagonal, register prefetch depth)
template <typename T, typename TA, int STAGE, int EPI, int FIN>
static hipError_t launch_ebe(bhip_ctx* c, EbeArgs<T, TA> a, bool trans, char typdia,
                             char typext) {
  const size_t lds = lds_bytes(c->L, sizeof(T));
  dim3 grid(c->L.nblocks), block(kThreads);
  const int diag = typdia == 'Q' ? DIAG_Q : (typdia == 'I' ? DIAG_ONE : DIAG_ZERO);
  // TYPEXT '0' is stored as zero symmetric coefficients
  const bool symq = typext == 'Q';
  const int mj = c->prefetch_j;  // pipeline depth 0..3 (triangles)
#define BHIP_LAUNCH(NDP, SQ, TR, DG, MJ)                                                      \
  hipLaunchKernelGGL((k_ebe<T, TA, NDP, SQ, TR, DG, STAGE, EPI, FIN, MJ>), grid, block, lds, \
                     c->stream, a)
#define BHIP_MJ(NDP, SQ, TR, DG)            \
  do {                                      \
    if (mj == 1)                            \
      BHIP_LAUNCH(NDP, SQ, TR, DG, 1);      \
    else if (mj == 2)                       \
      BHIP_LAUNCH(NDP, SQ, TR, DG, 2);      \
    else if (mj == 3)                       \
      BHIP_LAUNCH(NDP, SQ, TR, DG, 3);      \
    else                                    \
      BHIP_LAUNCH(NDP, SQ, TR, DG, 0);      \
  } while (0)
#define BHIP_DIAG(NDP, SQ, TR)              \
  do {                                      \
    if (diag == DIAG_Q)                     \
      BHIP_MJ(NDP, SQ, TR, DIAG_Q);         \
    else if (diag == DIAG_ONE)              \
      BHIP_MJ(NDP, SQ, TR, DIAG_ONE);       \
    else                                    \
      BHIP_MJ(NDP, SQ, TR, DIAG_ZERO);      \
  } while (0)
#define BHIP_PF1(NDP, SQ, TR, DG)           \
  do {                                      \
    if (mj > 0)                             \
      BHIP_LAUNCH(NDP, SQ, TR, DG, 1);      \
    else                                    \
      BHIP_LAUNCH(NDP, SQ, TR, DG, 0);      \
  } while (0)
#define BHIP_DIAG0(NDP, SQ, TR)              \
  do {                                       \
    if (diag == DIAG_Q)                      \
      BHIP_PF1(NDP, SQ, TR, DIAG_Q);         \
    else if (diag == DIAG_ONE)               \
      BHIP_PF1(NDP, SQ, TR, DIAG_ONE);       \
    else                                     \
      BHIP_PF1(NDP, SQ, TR, DIAG_ZERO);      \
  } while (0)
  if (c->L.ndp == 3) {
    if (!symq) BHIP_DIAG(3, false, false);
    else if (trans) BHIP_DIAG(3, true, true);
    else BHIP_DIAG(3, true, false);
  } else if (c->L.ndp == 2) {
    // segment storage: two-node "elements", one ('S') or two ('Q') terms
    if (!symq) BHIP_DIAG(2, false, false);
    else if (trans) BHIP_DIAG(2, true, true);
    else BHIP_DIAG(2, true, false);
  } else {
    // prisms: 15-30 coefficients per element, so at most one phase ahead
    if (!symq) BHIP_DIAG0(6, false, false);
    else if (trans) BHIP_DIAG0(6, true, true);
    else BHIP_DIAG0(6, true, false);
  }
#undef BHIP_DIAG0
#undef BHIP_PF1
#undef BHIP_DIAG
#undef BHIP_MJ
#undef BHIP_LAUNCH
  return hipGetLastError();
}

}  // namespace bhip

using namespace bhip;

// ---------------------------------------------------------------------------
// static scratch of bhip_solve_dev_noalloc: the per-workgroup partial sums
// (4 arrays of up to kScratchParts doubles) and the solver states of up to
// kScratchSys systems. Every GPU the library runs on holds its own copy until
// the process ends; all calls on one GPU share it.
// ---------------------------------------------------------------------------
constexpr int kScratchParts = 65536 + 64;  // node blocks (and node-kernel grid) + 64
constexpr int kScratchSys = 64;
constexpr size_t kScratchPartBytes = (size_t)4 * kScratchParts * sizeof(double);
constexpr size_t kScratchBytes = kScratchPartBytes + (size_t)kScratchSys * sizeof(SolveState);
static_assert(kScratchPartBytes % 256 == 0, "states start 256-byte aligned");
static_assert(sizeof(SolveState) % sizeof(double) == 0, "states are cleared as doubles");
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of `scratch` on device dev (the array's name is only a
// host-side handle), looked up once per device with that device current.
static void* scratch_addr(int dev) {
  static std::mutex mu;
  static void* addr[64] = {};
  if (dev < 0 || dev >= 64) return nullptr;
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[dev]) {
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess || cur != dev) return nullptr;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      fprintf(stderr, "libbiefhip: hipGetSymbolAddress(scratch) failed: %s\n",
              hipGetErrorString(e));
      return nullptr;
    }
    addr[dev] = p;
  }
  return addr[dev];
}

// ===========================================================================
// C ABI
// ===========================================================================
// Device side of a context whose layout (and, for ensembles, systems) is set.
static int ctx_setup(std::unique_ptr<bhip_ctx>& c, bhip_ctx** out) {
  const Layout& L = c->L;
  c->lds_f64 = lds_bytes(L, sizeof(double));
  c->lds_f32 = lds_bytes(L, sizeof(float));
  if (c->lds_f64 > 64 * 1024) {
    fprintf(stderr, "libbiefhip: block needs %zu B of LDS; use smaller blocks\n", c->lds_f64);
    return BHIP_EBADARG;
  }
  HIP_OK(hipGetDevice(&c->device));
  HIP_OK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  // the EBE product's many instantiations keep to the default 64 KiB of
  // dynamic LDS (checked above); only the prism element kernels raise their
  // limit, per launch, for their node records
  const int n = L.npoin;
  c->grid_nodes = grid_for(n);
  HIP_OK(c->node_start.alloc(L.node_start.size()));
  HIP_OK(c->halo_start.alloc(L.halo_start.size()));
  HIP_OK(c->halo.alloc(std::max<size_t>(1, L.halo.size())));
  HIP_OK(c->color_off.alloc(L.color_off.size()));
  HIP_OK(c->lid.alloc(L.lid.size()));
  HIP_OK(c->slot_elem.alloc(L.slot_elem.size()));
  HIP_OK(c->perm.alloc(L.perm.size()));
  HIP_OK(c->first_slot.alloc(L.nelem));
  std::vector<int32_t> first(L.nelem, -1);
  for (int s = 0; s < L.nslot; ++s)
    if (first[L.slot_elem[s]] < 0) first[L.slot_elem[s]] = s;
  auto up = [&](void* dst, const void* src, size_t bytes) {
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream);
  };
  HIP_OK(up(c->node_start.p, L.node_start.data(), L.node_start.size() * 4));
  HIP_OK(up(c->halo_start.p, L.halo_start.data(), L.halo_start.size() * 4));
  if (!L.halo.empty()) HIP_OK(up(c->halo.p, L.halo.data(), L.halo.size() * 4));
  HIP_OK(up(c->color_off.p, L.color_off.data(), L.color_off.size() * 4));
  HIP_OK(up(c->lid.p, L.lid.data(), L.lid.size() * 2));
  HIP_OK(up(c->slot_elem.p, L.slot_elem.data(), L.slot_elem.size() * 4));
  HIP_OK(up(c->perm.p, L.perm.data(), L.perm.size() * 4));
  HIP_OK(up(c->first_slot.p, first.data(), first.size() * 4));
  // work memory
  for (auto* v : {&c->vx, &c->vb, &c->vr, &c->vd0, &c->vd1, &c->vad, &c->vdz, &c->vdinv,
                  &c->vio0, &c->vio1})
    HIP_OK(v->alloc(n));
  c->max_part = std::max(L.nblocks, c->grid_nodes) + 64;
  HIP_OK(c->partial.alloc((size_t)4 * c->max_part));
  hipLaunchKernelGGL(k_fill<double>, dim3(grid_for(4 * c->max_part)), dim3(kThreads), 0,
                     c->stream, c->partial.p, 0.0, (long)4 * c->max_part);
  HIP_OK(hipGetLastError());
  if (const char* ev = getenv("BHIP_NO_GRAPH")) c->use_graphs = !(ev[0] == '1');
  // reliable updates in the two-kernel inner CG; BHIP_RELIABLE=0 restarts
  // every inner solve from D = R instead
  c->reliable = true;
  if (const char* ev = getenv("BHIP_RELIABLE")) c->reliable = ev[0] != '0';
  c->prefetch_j = 1;
  if (const char* ev = getenv("BHIP_PREFETCH")) {
    int v = atoi(ev);
    if (v >= 0 && v <= 3) c->prefetch_j = v;
  }
  HIP_OK(c->st.alloc(c->nsys));
  HIP_OK(hipHostMalloc((void**)&c->st_host, sizeof(SolveState) * c->nsys, hipHostMallocDefault));
  std::memset(c->st_host, 0, sizeof(SolveState) * c->nsys);
  HIP_OK(up(c->st.p, c->st_host, sizeof(SolveState) * c->nsys));
  c->part_p = c->partial.p;
  c->st_p = c->st.p;
  // the static scratch of bhip_solve_dev_noalloc, located once per device
  if (scratch_addr(c->device) == nullptr) return BHIP_EDEVICE;
  c->sys_info.assign(c->nsys, bhip_solve_info());
  if (c->nsys > 1) {
    std::vector<int32_t> bs(L.nblocks);
    for (int sy = 0; sy < c->nsys; ++sy)
      for (int b = c->sys_blocks_h[sy]; b < c->sys_blocks_h[sy + 1]; ++b) bs[b] = sy;
    HIP_OK(c->block_sys.alloc(L.nblocks));
    HIP_OK(c->sys_blocks.alloc(c->nsys + 1));
    HIP_OK(up(c->block_sys.p, bs.data(), bs.size() * 4));
    HIP_OK(up(c->sys_blocks.p, c->sys_blocks_h.data(), c->sys_blocks_h.size() * 4));
  }
  hipLaunchKernelGGL(k_fill<double>, dim3(c->grid_nodes), dim3(kThreads), 0, c->stream,
                     c->vdz.p, 0.0, (long)n);
  HIP_OK(hipGetLastError());
  HIP_OK(hipStreamSynchronize(c->stream));
  *out = c.release();
  return BHIP_OK;
}

extern "C" {

int bhip_ctx_create(int npoin, int nelem, int ndp, const int32_t* ikle, const double* x,
                    const double* y, int nplan, int block_nodes, bhip_ctx** out) {
  if (!out || !ikle) return BHIP_EBADARG;
  *out = nullptr;
  std::unique_ptr<bhip_ctx> c(new bhip_ctx());
  std::string err = build_layout(npoin, nelem, ndp, ikle, x, y, nplan, block_nodes, c->L);
  if (!err.empty()) {
    fprintf(stderr, "libbiefhip: bhip_ctx_create: %s\n", err.c_str());
    return BHIP_EBADARG;
  }
  return ctx_setup(c, out);
}

int bhip_batch_create(int nsys, const int* npoin, const int* nelem, int ndp,
                      const int32_t* ikle, const double* x, const double* y, int block_nodes,
                      bhip_ctx** out) {
  if (!out || !ikle || !npoin || !nelem || nsys < 1 || (ndp != 2 && ndp != 3)) return BHIP_EBADARG;
  *out = nullptr;
  std::unique_ptr<bhip_ctx> c(new bhip_ctx());
  std::vector<int32_t> ikle_union;
  std::string err = build_layout_multi(nsys, npoin, nelem, ndp, ikle, x, y, block_nodes, c->L,
                                       c->sys_blocks_h, ikle_union);
  if (!err.empty()) {
    fprintf(stderr, "libbiefhip: bhip_batch_create: %s\n", err.c_str());
    return BHIP_EBADARG;
  }
  c->nsys = nsys;
  c->sys_node_off.assign(nsys + 1, 0);
  c->sys_elem_off.assign(nsys + 1, 0);
  for (int s = 0; s < nsys; ++s) {
    c->sys_node_off[s + 1] = c->sys_node_off[s] + npoin[s];
    c->sys_elem_off[s + 1] = c->sys_elem_off[s] + nelem[s];
  }
  return ctx_setup(c, out);
}

int bhip_batch_info(const bhip_ctx* c, int nsys, bhip_solve_info* infos) {
  if (!c || !infos || nsys != c->nsys) return BHIP_EBADARG;
  for (int s = 0; s < nsys; ++s) infos[s] = c->sys_info[s];
  return BHIP_OK;
}

int bhip_batch_latency(const bhip_ctx* c, int nsys, double* ms) {
  if (!c || !ms || nsys != c->nsys || (int)c->sys_latency_ms.size() != nsys) return BHIP_EBADARG;
  for (int s = 0; s < nsys; ++s) ms[s] = c->sys_latency_ms[s];
  return BHIP_OK;
}


int bhip_ctx_destroy(bhip_ctx* c) {
  if (!c) return BHIP_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto* v : {&c->node_start, &c->halo_start, &c->halo, &c->color_off, &c->slot_elem,
                  &c->perm, &c->first_slot})
    v->release();
  c->lid.release();
  for (auto* v : {&c->bicg_sc, &c->bicg_part, &c->gm, &c->gm_scal, &c->da_saved, &c->xa_saved,
                  &c->da, &c->xa, &c->vx, &c->vb, &c->vr, &c->vd0, &c->vd1, &c->vad, &c->vdz,
                  &c->vdinv, &c->vio0, &c->vio1, &c->vx2, &c->partial})
    v->release();
  for (auto* v : {&c->da32, &c->xa32, &c->fr, &c->fd0, &c->fd1, &c->fad, &c->fx0, &c->fdz,
                  &c->gm32})
    v->release();
  c->geo.release();
  c->st.release();
  for (auto* v : {&c->wv_geo, &c->wv_w, &c->wv_vol, &c->wv_h, &c->wv_zb, &c->wv_f, &c->wv_vs,
                  &c->wv_c, &c->wv_usx, &c->wv_usy, &c->wv_fb, &c->wv_da, &c->wv_sbx,
                  &c->wv_sby, &c->wv_svs, &c->wv_seb})
    v->release();
  for (auto& gr : c->graphs)
    if (gr.exec) (void)hipGraphExecDestroy(gr.exec);
  for (auto& sg : c->solve_graphs)
    if (sg.exec) (void)hipGraphExecDestroy(sg.exec);
  if (c->st_host) (void)hipHostFree(c->st_host);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return BHIP_OK;
}

void* bhip_ctx_stream(bhip_ctx* c) { return c ? (void*)c->stream : nullptr; }

int bhip_layout_check(int npoin, int nelem, int ndp, const int32_t* ikle, const double* x,
                      const double* y, int nplan, int block_nodes, int64_t* stats) {
  Layout L;
  std::string err = build_layout(npoin, nelem, ndp, ikle, x, y, nplan, block_nodes, L);
  if (!err.empty()) {
    fprintf(stderr, "libbiefhip: bhip_layout_check: %s\n", err.c_str());
    return BHIP_EBADARG;
  }
  if (stats) {
    stats[0] = L.nblocks;
    stats[1] = L.max_colors;
    stats[2] = L.nslot;
    stats[3] = L.halo_total;
    stats[4] = L.max_local;
    stats[5] = L.max_owned;
    stats[6] = (int64_t)lds_bytes(L, sizeof(double));
  }
  return BHIP_OK;
}

int bhip_ctx_stats(const bhip_ctx* c, int* nblocks, int* max_colors, int64_t* elem_instances,
                   int64_t* halo_nodes) {
  if (!c) return BHIP_EBADARG;
  if (nblocks) *nblocks = c->L.nblocks;
  if (max_colors) *max_colors = c->L.max_colors;
  if (elem_instances) *elem_instances = c->L.nslot;
  if (halo_nodes) *halo_nodes = c->L.halo_total;
  return BHIP_OK;
}

static int matrix_set_common(bhip_ctx* c, char typdia, char typext, const double* da,
                             const double* xa, hipMemcpyKind kind) {
  if (!c) return BHIP_EBADARG;
  if (typdia != 'Q' && typdia != 'I' && typdia != '0') return BHIP_EBADARG;
  if (typext != 'Q' && typext != 'S' && typext != '0') return BHIP_EBADARG;
  const Layout& L = c->L;
  const int npair = L.ndp * (L.ndp - 1) / 2;
  const int ncoef = typext == 'Q' ? 2 * npair : (typext == 'S' ? npair : 0);
  if (ncoef > 0 && !xa) return BHIP_EBADARG;
  HIP_OK(c->da.alloc(L.npoin));
  // coefficient storage always holds the non-symmetric width, so a 'S'
  // matrix can be turned 'Q' later without reallocation
  HIP_OK(c->xa.alloc((size_t)2 * npair * L.nslot_pad));
  if (da) {
    // DA in internal order
    double* tmp = c->vio0.p;
    if (kind == hipMemcpyHostToDevice)
      HIP_OK(to_device(c, tmp, da, sizeof(double) * L.npoin));
    else
      HIP_OK(hipMemcpyAsync(tmp, da, sizeof(double) * L.npoin, kind, c->stream));
    hipLaunchKernelGGL(k_gather<double>, dim3(c->grid_nodes), dim3(kThreads), 0, c->stream,
                       c->da.p, tmp, c->perm.p, L.npoin);
  } else {
    hipLaunchKernelGGL(k_fill<double>, dim3(c->grid_nodes), dim3(kThreads), 0, c->stream,
                       c->da.p, 0.0, (long)L.npoin);
  }
  HIP_OK(hipGetLastError());
  if (ncoef > 0) {
    double* src = nullptr;
    if (kind == hipMemcpyHostToDevice) {
      HIP_OK(c->h_xa.alloc((size_t)ncoef * L.nelem));
      HIP_OK(to_device(c, c->h_xa.p, xa, sizeof(double) * ncoef * L.nelem));
      src = c->h_xa.p;
    } else {
      src = const_cast<double*>(xa);
    }
    hipLaunchKernelGGL(k_matrix_to_slots, dim3(grid_for(L.nslot * ncoef)), dim3(kThreads), 0,
                       c->stream, c->xa.p, src, c->slot_elem.p, L.nslot, L.nslot_pad, L.nelem,
                       ncoef);
    HIP_OK(hipGetLastError());
  } else {
    hipLaunchKernelGGL(k_fill<double>, dim3(grid_for(L.nslot_pad)), dim3(kThreads), 0,
                       c->stream, c->xa.p, 0.0, (long)npair * L.nslot_pad);
    HIP_OK(hipGetLastError());
  }
  c->typdia = typdia;
  c->typext = typext;
  c->ncoef = ncoef;
  c->have_matrix = true;
  // host arrays are read before returning; device arrays asynchronously,
  // in stream order on the context stream
  if (kind == hipMemcpyHostToDevice) HIP_OK(hipStreamSynchronize(c->stream));
  return BHIP_OK;
}

int bhip_matrix_set(bhip_ctx* c, char typdia, char typext, const double* da, const double* xa) {
  return matrix_set_common(c, typdia, typext, da, xa, hipMemcpyHostToDevice);
}
int bhip_matrix_set_dev(bhip_ctx* c, char typdia, char typext, const double* da,
                        const double* xa) {
  return matrix_set_common(c, typdia, typext, da, xa, hipMemcpyDeviceToDevice);
}

int bhip_matrix_save(bhip_ctx* c) {
  if (!c || !c->have_matrix) return BHIP_EBADARG;
  const Layout& L = c->L;
  const int npair = L.ndp * (L.ndp - 1) / 2;
  HIP_OK(c->da_saved.alloc(L.npoin));
  HIP_OK(c->xa_saved.alloc((size_t)2 * npair * L.nslot_pad));
  hipLaunchKernelGGL(k_copy, dim3(1024), dim3(kThreads), 0, c->stream, c->da_saved.p, c->da.p,
                     (long)L.npoin);
  hipLaunchKernelGGL(k_copy, dim3(1024), dim3(kThreads), 0, c->stream, c->xa_saved.p, c->xa.p,
                     (long)std::max(c->ncoef, npair) * L.nslot_pad);
  HIP_OK(hipGetLastError());
  c->typdia_saved = c->typdia;
  c->typext_saved = c->typext;
  c->ncoef_saved = c->ncoef;
  c->have_saved = true;
  HIP_OK(hipStreamSynchronize(c->stream));
  return BHIP_OK;
}

int bhip_matrix_restore(bhip_ctx* c) {
  if (!c || !c->have_saved) return BHIP_EBADARG;
  const Layout& L = c->L;
  const int npair = L.ndp * (L.ndp - 1) / 2;
  hipLaunchKernelGGL(k_copy, dim3(1024), dim3(kThreads), 0, c->stream, c->da.p, c->da_saved.p,
                     (long)L.npoin);
  hipLaunchKernelGGL(k_copy, dim3(1024), dim3(kThreads), 0, c->stream, c->xa.p, c->xa_saved.p,
                     (long)std::max(c->ncoef_saved, npair) * L.nslot_pad);
  HIP_OK(hipGetLastError());
  c->typdia = c->typdia_saved;
  c->typext = c->typext_saved;
  c->ncoef = c->ncoef_saved;
  c->have_matrix = true;
  return BHIP_OK;
}

int bhip_matrix_info(const bhip_ctx* c, char* typdia, char* typext, int* ncoef) {
  if (!c) return BHIP_EBADARG;
  const bool m = c->have_matrix;
  if (typdia) *typdia = m ? c->typdia : '0';
  if (typext) *typext = m ? c->typext : '0';
  if (ncoef) *ncoef = m ? c->ncoef : 0;
  return BHIP_OK;
}

int bhip_matrix_get(bhip_ctx* c, char* typdia, double* da, double* xa) {
  if (!c || !c->have_matrix) return BHIP_EBADARG;
  const Layout& L = c->L;
  if (typdia) *typdia = c->typdia;
  if (da) {
    hipLaunchKernelGGL(k_scatter<double>, dim3(c->grid_nodes), dim3(kThreads), 0, c->stream,
                       c->vio0.p, c->da.p, c->perm.p, L.npoin);
    HIP_OK(hipGetLastError());
    HIP_OK(to_host(c, da, c->vio0.p, sizeof(double) * L.npoin));
  }
  if (xa && c->ncoef > 0) {
    HIP_OK(c->h_xa.alloc((size_t)c->ncoef * L.nelem));
    hipLaunchKernelGGL(k_matrix_from_slots, dim3(grid_for(L.nelem * c->ncoef)), dim3(kThreads),
                       0, c->stream, c->h_xa.p, c->xa.p, c->first_slot.p, L.nelem, L.nslot_pad,
                       c->ncoef);
    HIP_OK(hipGetLastError());
    HIP_OK(to_host(c, xa, c->h_xa.p, sizeof(double) * c->ncoef * L.nelem));
  }
  HIP_OK(hipStreamSynchronize(c->stream));
  return BHIP_OK;
}

// y (caller order, device) -> vio0 (internal); result vio1 (internal) ->
// x (caller order, device)
static int matvec_internal(bhip_ctx* c, int op, double cst, const double* y_in,
                           const double* x_in, double* x_out) {
  EbeArgs<double, double> a = base_args<double, double>(c, c->xa.p, c->da.p);
  a.y = y_in;
  a.xin = x_in;
  a.out = x_out;
  a.op = op;
  a.c = cst;
  const bool trans = op >= BHIP_OP_TAY;
  HIP_OK((launch_ebe<double, double, STAGE_PLAIN, EPI_OP, FIN_NONE>(c, a, trans, c->typdia,
                                                                     c->typext)));
  return BHIP_OK;
}

int bhip_matvec_dev(bhip_ctx* c, int op, double cst, const double* y, double* x) {
  if (!c || !c->have_matrix || op < 0 || op > 10 || !y || !x) return BHIP_EBADARG;
  const Layout& L = c->L;
  const int g = c->grid_nodes;
  hipLaunchKernelGGL(k_gather<double>, dim3(g), dim3(kThreads), 0, c->stream, c->vio0.p, y,
                     c->perm.p, L.npoin);
  hipLaunchKernelGGL(k_gather<double>, dim3(g), dim3(kThreads), 0, c->stream, c->vio1.p, x,
                     c->perm.p, L.npoin);
  HIP_OK(hipGetLastError());
  int rc = matvec_internal(c, op, cst, c->vio0.p, c->vio1.p, c->vr.p);
  if (rc) return rc;
  hipLaunchKernelGGL(k_scatter<double>, dim3(g), dim3(kThreads), 0, c->stream, x, c->vr.p,
                     c->perm.p, L.npoin);
  HIP_OK(hipGetLastError());
  return BHIP_OK;
}

int bhip_matvec(bhip_ctx* c, int op, double cst, const double* y, double* x) {
  if (!c || !c->have_matrix || !y || !x) return BHIP_EBADARG;
  const Layout& L = c->L;
  const size_t bytes = sizeof(double) * L.npoin;
  HIP_OK(c->h_b.alloc(L.npoin));
  HIP_OK(c->h_x.alloc(L.npoin));
  HIP_OK(to_device(c, c->h_b.p, y, bytes));
  HIP_OK(to_device(c, c->h_x.p, x, bytes));
  int rc = bhip_matvec_dev(c, op, cst, c->h_b.p, c->h_x.p);
  if (rc == BHIP_OK) HIP_OK(to_host(c, x, c->h_x.p, bytes));
  (void)hipStreamSynchronize(c->stream);
  return rc;
}

}  // extern "C"

// ===========================================================================
// SOLVE
// ===========================================================================
namespace bhip {

static int read_state(bhip_ctx* c) {
  HIP_OK(hipMemcpyAsync(c->st_host, c->st_p, sizeof(SolveState) * c->nsys, hipMemcpyDeviceToHost,
                        c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  return BHIP_OK;
}

static int write_state(bhip_ctx* c) {
  HIP_OK(hipMemcpyAsync(c->st_p, c->st_host, sizeof(SolveState) * c->nsys, hipMemcpyHostToDevice,
                        c->stream));
  return BHIP_OK;
}

static double* part_u(bhip_ctx* c, int p) { return c->part_p + (size_t)p * c->max_part; }
static double* part_e(bhip_ctx* c, int p) { return c->part_p + (size_t)(2 + p) * c->max_part; }

// host copy of the state -> device (the device never writes a field the
// host sets here while no kernel is in flight)
static void init_cg_state(SolveState* sh, double eps2xl, int use_rr0, int nitmax) {
  sh->eps2xl = eps2xl;
  sh->use_rr0 = use_rr0;
  // the reference always runs iteration 1 before testing NITMAX (gracjg.F:406)
  sh->nitmax = std::max(nitmax, 1);
  sh->m[0] = sh->m[1] = -1;
  sh->status[0] = sh->status[1] = ST_RUNNING;
  sh->rr[0] = sh->rr[1] = 1.0;
  sh->done = 0;
  sh->done1 = 0;
  sh->stop = 0;
  sh->final_flag = 0;
  sh->final_m = 0;
  sh->final_rr = 0.0;
}

// One batch of `iters` CG iterations (each U then E; an even count keeps the
// direction ping-pong aligned). All kernels exit at once when the solve is
// done, so a batch can run past convergence at no cost but launches.
// the X/R update kernel of iteration parity p (one workgroup per node block)
template <typename T, typename TX>
static hipError_t launch_update(bhip_ctx* c, TX* x, T* r, const T* d, const T* ad, int p) {
  hipLaunchKernelGGL((k_cg_update<T, TX>), dim3(c->L.nblocks), dim3(kThreads), 0, c->stream, x,
                     r, d, ad, c->node_start.p, c->L.nblocks, sys_ptr(c->block_sys),
                     sys_ptr(c->sys_blocks), part_e(c, 1 - p), part_u(c, p), c->st_p, p);
  return hipGetLastError();
}

template <typename T, typename TA, typename TX>
static hipError_t enqueue_cg_batch(bhip_ctx* c, int iters, TX* x, T* r, T* d0, T* d1, T* ad,
                                   const TA* xa, const TA* da) {
  for (int k = 0; k < iters; ++k) {
    const int p = (k & 1) ? 0 : 1;  // iterations 1, 2, 3, ... of the batch
    T* dprev = (p == 1) ? d0 : d1;
    T* dnext = (p == 1) ? d1 : d0;
    hipError_t e = launch_update<T, TX>(c, x, r, dprev, ad, p);
    if (e != hipSuccess) return e;
    EbeArgs<T, TA> a = base_args<T, TA>(c, xa, da);
    a.r = r;
    a.dold = dprev;
    a.dnew = dnext;
    a.out = ad;
    a.partial = part_e(c, p);
    a.part_in = part_u(c, p);
    a.n_part_in = c->L.nblocks;
    a.parity = p;
    e = launch_ebe<T, TA, STAGE_CGDIR, EPI_CGAD, FIN_CG>(c, a, false, c->typdia, c->typext);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// The graph (or plain launches) of one batch of kBatch CG iterations.
constexpr int kBatch = 8;
static int even_up(int n) { return n + (n & 1); }

// A captured batch of CG iterations, cached per (solution buffer, kernel
// family, TYPDIA/TYPEXT specialisation).
template <typename F>
static int cached_graph(bhip_ctx* c, const void* key, int prec, F&& enqueue,
                        hipGraphExec_t* out) {
  *out = nullptr;
  if (!c->use_graphs) return BHIP_OK;
  int slot = -1;
  const int tag = ((int)(unsigned char)c->typdia << 8) | (int)(unsigned char)c->typext;
  for (int i = 0; i < 4; ++i)
    if (c->graphs[i].exec && c->graphs[i].key == key && c->graphs[i].prec == prec &&
        c->graphs[i].tag == tag)
      slot = i;
  if (slot < 0) {
    for (int i = 0; i < 4 && slot < 0; ++i)
      if (!c->graphs[i].exec) slot = i;
    if (slot < 0) {
      (void)hipGraphExecDestroy(c->graphs[3].exec);
      c->graphs[3].exec = nullptr;
      slot = 3;
    }
    hipGraph_t graph;
    HIP_OK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = enqueue();
    hipError_t e2 = hipStreamEndCapture(c->stream, &graph);
    HIP_OK(e);
    HIP_OK(e2);
    HIP_OK(hipGraphInstantiate(&c->graphs[slot].exec, graph, nullptr, nullptr, 0));
    (void)hipGraphDestroy(graph);
    c->graphs[slot].key = key;
    c->graphs[slot].prec = prec;
    c->graphs[slot].tag = tag;
  }
  *out = c->graphs[slot].exec;
  return BHIP_OK;
}

template <typename T, typename TA, typename TX>
static int cg_batch_exec(bhip_ctx* c, int prec, TX* x, T* r, T* d0, T* d1, T* ad,
                         const TA* xa, const TA* da, hipGraphExec_t* out) {
  return cached_graph(c, (const void*)x, prec, [&]() {
    return enqueue_cg_batch<T, TA, TX>(c, kBatch, x, r, d0, d1, ad, xa, da);
  }, out);
}

// Runs `plan` (a function queuing kernels only) on the context stream: as a
// graph captured on first use for this key and replayed afterwards, or
// directly when graphs are off.
template <typename F>
static int run_planned(bhip_ctx* c, const uint64_t (&key)[8], F&& plan) {
  if (!c->use_graphs) return plan(false);
  hipGraphExec_t exec = nullptr;
  for (auto& sg : c->solve_graphs)
    if (std::equal(std::begin(key), std::end(key), std::begin(sg.key))) exec = sg.exec;
  if (!exec) {
    if (c->solve_graphs.size() >= 8) {
      (void)hipGraphExecDestroy(c->solve_graphs.front().exec);
      c->solve_graphs.erase(c->solve_graphs.begin());
    }
    hipGraph_t graph;
    HIP_OK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    const int rc = plan(true);
    hipError_t e2 = hipStreamEndCapture(c->stream, &graph);
    if (rc) return rc;
    HIP_OK(e2);
    HIP_OK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    (void)hipGraphDestroy(graph);
    bhip_ctx::SolveGraph sg;
    std::copy(std::begin(key), std::end(key), std::begin(sg.key));
    sg.exec = exec;
    c->solve_graphs.push_back(sg);
  }
  HIP_OK(hipGraphLaunch(exec, c->stream));
  return BHIP_OK;
}

static uint64_t dbits(double v) {
  uint64_t u;
  std::memcpy(&u, &v, sizeof(u));
  return u;
}

static uint64_t matrix_tag(const bhip_ctx* c) {
  return ((uint64_t)(unsigned char)c->typdia << 8) | (uint64_t)(unsigned char)c->typext;
}

template <typename T, typename TA, typename TX>
static int enqueue_batches(bhip_ctx* c, hipGraphExec_t exec, int nb, TX* x, T* r, T* d0, T* d1,
                           T* ad, const TA* xa, const TA* da) {
  for (int k = 0; k < nb; ++k) {
    if (exec)
      HIP_OK(hipGraphLaunch(exec, c->stream));
    else
      HIP_OK((enqueue_cg_batch<T, TA, TX>(c, kBatch, x, r, d0, d1, ad, xa, da)));
  }
  return BHIP_OK;
}

// E_0 of a CG solve: D = R, AD, RO (gracjg.F:248-268); R's partial sums of
// ||R||^2 are in part_u[0] (n_in of them)
template <typename T, typename TA>
static hipError_t enqueue_cg_start(bhip_ctx* c, T* r, T* d0, T* ad, const T* dz, const TA* xa,
                                   const TA* da, int n_in, T* xzero) {
  EbeArgs<T, TA> a = base_args<T, TA>(c, xa, da);
  a.r = r;
  a.dold = dz;
  a.dnew = d0;
  a.out = ad;
  a.partial = part_e(c, 0);
  a.part_in = part_u(c, 0);
  a.n_part_in = n_in;
  a.parity = 0;
  // the first kernel of a mixed inner solve: only for systems starting one,
  // and it zeroes their inner solution
  a.skip_unless_fresh = xzero != nullptr;
  a.xzero = xzero;
  return launch_ebe<T, TA, STAGE_CGDIR, EPI_CGAD, FIN_CG>(c, a, false, c->typdia, c->typext);
}

static hipError_t enqueue_bb_setup(bhip_ctx* c, const double* b, double eps, int nitmax,
                                   int mixed, double inner_eps, int inner_nitmax,
                                   int reliable = 0) {
  hipLaunchKernelGGL(k_dot_blocks, dim3(c->L.nblocks), dim3(kThreads), 0, c->stream, b, b,
                     c->node_start.p, part_u(c, 1));
  hipLaunchKernelGGL(k_cg_setup, dim3(c->nsys), dim3(kThreads), 0, c->stream, part_u(c, 1),
                     c->L.nblocks, sys_ptr(c->sys_blocks), eps, nitmax, mixed, inner_eps,
                     inner_nitmax, reliable, c->st_p);
  return hipGetLastError();
}

static int ensure_f32(bhip_ctx* c) {
  const Layout& L = c->L;
  const int n = L.npoin;
  const bool fresh = !c->fdz.p || c->fdz.n < (size_t)n;
  for (auto* v : {&c->fr, &c->fd0, &c->fd1, &c->fad, &c->fx0, &c->fdz}) HIP_OK(v->alloc(n));
  HIP_OK(c->da32.alloc(n));
  const int npair = L.ndp * (L.ndp - 1) / 2;
  HIP_OK(c->xa32.alloc((size_t)2 * npair * L.nslot_pad));
  // a fresh inner start multiplies fdz by BETA = 0, a continued one reads it
  // after k_inner_dir has written a direction there: it only ever needs to
  // hold finite values, so it is zeroed once, when allocated
  if (fresh) {
    hipLaunchKernelGGL(k_fill<float>, dim3(c->grid_nodes), dim3(kThreads), 0, c->stream,
                       c->fdz.p, 0.0f, (long)n);
    HIP_OK(hipGetLastError());
  }
  return BHIP_OK;
}

// device clock rate (kHz) of wall_clock64(), for the per-system latencies
static double clock_khz(bhip_ctx* c) {
  int khz = 0;
  if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->device) != hipSuccess ||
      khz <= 0)
    return 0.0;
  return (double)khz;
}

// per-system outcome of a CG solve (FP64 or mixed) from the device states;
// fills c->sys_info and the ensemble summary in *info
static int collect_infos(bhip_ctx* c, double* x, bool mixed, bhip_solve_info* info) {
  SolveState* sh = c->st_host;
  const int g = c->grid_nodes;
  const double khz = clock_khz(c);
  c->sys_latency_ms.assign(c->nsys, 0.0);
  int maxm = 0, nconv = 0, nzero = 0, maxout = 0, maxin = 0;
  double maxrel = 0.0;
  for (int sy = 0; sy < c->nsys; ++sy) {
    const SolveState& q = sh[sy];
    bhip_solve_info& si = c->sys_info[sy];
    std::memset(&si, 0, sizeof(si));
    si.relat = q.xl_raw >= 1.0;
    if (khz > 0 && q.t_done >= q.t_start) c->sys_latency_ms[sy] = (q.t_done - q.t_start) / khz;
    if (q.zero_rhs) {
      // X = 0 (gracjg.F:194-202) on this system's nodes
      const int b0 = c->nsys > 1 ? c->sys_blocks_h[sy] : 0;
      const int b1 = c->nsys > 1 ? c->sys_blocks_h[sy + 1] : c->L.nblocks;
      const int n0 = c->L.node_start[b0], n1 = c->L.node_start[b1];
      hipLaunchKernelGGL(k_fill<double>, dim3(g), dim3(kThreads), 0, c->stream, x + n0, 0.0,
                         (long)(n1 - n0));
      HIP_OK(hipGetLastError());
      si.converged = 1;
      ++nzero;
      ++nconv;
      continue;
    }
    if (mixed) {
      si.niter = q.inner_total;
      si.inner_iters = q.inner_total;
      si.outer_iters = q.outer_count;
      si.relres = std::sqrt(q.rr_outer / q.xl);
      si.converged = q.outer_status == ST_CONVERGED;
    } else {
      si.niter = q.final_m;
      si.relres = std::sqrt(q.final_rr / q.xl);
      si.converged = q.done == ST_CONVERGED;
    }
    nconv += si.converged;
    maxm = std::max(maxm, si.niter);
    maxin = std::max(maxin, si.inner_iters);
    maxout = std::max(maxout, si.outer_iters);
    maxrel = std::max(maxrel, si.relres);
  }
  if (nzero == c->nsys) {
    *info = c->sys_info[0];
    return BHIP_ZERO_RHS;
  }
  if (c->nsys == 1) {
    *info = c->sys_info[0];
  } else {
    // the ensemble: largest counts and precision, converged if all did
    info->niter = maxm;
    info->inner_iters = maxin;
    info->outer_iters = maxout;
    info->relres = maxrel;
    info->relat = c->sys_info[0].relat;
    info->converged = nconv == c->nsys;
  }
  return info->converged ? BHIP_OK : BHIP_MAXITER;
}

// GRACJG (gracjg.F:184-420) on the device-resident preconditioned system,
// FP64. The whole solve is queued (set-up, R = A X - B, E_0, as many
// iteration batches as the previous solve needed); the host then waits once
// and only queues more batches if the solve is not finished.
static int cg_fp64(bhip_ctx* c, double* x, const double* b, double eps, int nitmax,
                   bhip_solve_info* info, const std::function<int()>& tail) {
  const int n = c->L.npoin;
  const int g = c->grid_nodes;
  SolveState* sh = c->st_host;
  hipGraphExec_t exec;
  int rc = cg_batch_exec<double, double, double>(c, BHIP_PREC_FP64, x, c->vr.p, c->vd0.p,
                                                 c->vd1.p, c->vad.p, c->xa.p, c->da.p, &exec);
  if (rc) return rc;
  int& hint = c->iter_hint[0];
  int nb = 1;  // batches of the first extra round (doubling up to 8)
  // iterations queued up front: as many as the previous solve needed (each
  // launched iteration past convergence still costs two launches), rounded
  // up to even so that later batches start on the same parity
  const int npl = hint > 0 ? even_up(hint) : kBatch;
  // set-up, R = A X - B (gracjg.F:213-215) with the partial sums of RMRM,
  // E_0, and the iterations the previous solve needed
  auto plan = [&](bool capturing) -> int {
    HIP_OK(enqueue_bb_setup(c, b, eps, nitmax, 0, 0.0, 0));
    EbeArgs<double, double> a = base_args<double, double>(c, c->xa.p, c->da.p);
    a.y = x;
    a.b = b;
    a.out = c->vr.p;
    a.partial = part_u(c, 0);
    HIP_OK((launch_ebe<double, double, STAGE_PLAIN, EPI_RESID, FIN_PARTIAL>(
        c, a, false, c->typdia, c->typext)));
    HIP_OK((enqueue_cg_start<double, double>(c, c->vr.p, c->vd0.p, c->vad.p, c->vdz.p, c->xa.p,
                                             c->da.p, c->L.nblocks, nullptr)));
    (void)capturing;
    HIP_OK((enqueue_cg_batch<double, double, double>(c, npl, x, c->vr.p, c->vd0.p, c->vd1.p,
                                                     c->vad.p, c->xa.p, c->da.p)));
    return BHIP_OK;
  };
  const uint64_t key[8] = {0, (uint64_t)npl, matrix_tag(c), dbits(eps), (uint64_t)nitmax,
                           (uint64_t)(uintptr_t)x, 0, 0};
  if ((rc = run_planned(c, key, plan))) return rc;
  // every system of an ensemble runs until its own stop; batches go on
  // while one of them has not stopped. The caller's results are written
  // before each wait, so a finished solve needs no launch after it
  for (;;) {
    if ((rc = tail())) return rc;
    if (read_state(c)) return BHIP_EDEVICE;
    bool all = true;
    for (int sy = 0; sy < c->nsys; ++sy) all = all && sh[sy].done != 0;
    if (all) break;
    rc = enqueue_batches<double, double, double>(c, exec, nb, x, c->vr.p, c->vd0.p, c->vd1.p,
                                                 c->vad.p, c->xa.p, c->da.p);
    nb = std::min(nb * 2, 8);
    if (rc) return rc;
  }
  int maxm = 0;
  for (int sy = 0; sy < c->nsys; ++sy)
    if (!sh[sy].zero_rhs) maxm = std::max(maxm, sh[sy].final_m);
  if (maxm > 0) hint = maxm;
  return collect_infos(c, x, false, info);
}


// Mixed precision CG: FP64 true residual r = A'x - b' and FP64 update
// x <- x - delta, with delta from an FP32 EBE CG on the FP32 copy of A'
// (inner stop ||R||^2 <= inner_eps^2 ||r||^2). Same stop rule as GRACJG on the
// preconditioned system, evaluated with the FP64 residual; NITMAX caps the
// inner iterations of the whole solve. The outer loop runs on the device,
// per system of an ensemble: each outer step (x -= delta fused into the FP64
// residual product, then k_outer_check) acts on the systems whose inner
// solve has finished, each inner start on the systems an outer test has just
// restarted, and the inner CG kernels on the systems still iterating. So
// the expected number of outer and inner steps is queued at once and the
// host waits once, then queues more rounds while a system has not stopped.
static int cg_mixed(bhip_ctx* c, double* x, const double* b, double eps, int nitmax,
                    double inner_eps, int inner_nitmax, bool xa32_ready,
                    bhip_solve_info* info, const std::function<int()>& tail) {
  const Layout& L = c->L;
  const int n = L.npoin;
  const int g = c->grid_nodes;
  SolveState* sh = c->st_host;
  if (ensure_f32(c)) return BHIP_EDEVICE;
  HIP_OK(c->vx2.alloc(n));
  const int npair = L.ndp * (L.ndp - 1) / 2;
  const long ncf = (long)(c->typext == 'Q' ? 2 * npair : npair) * L.nslot_pad;
  float* delta = c->fx0.p;
  hipGraphExec_t exec;
  int rc = cg_batch_exec<float, float, float>(c, BHIP_PREC_MIXED, delta, c->fr.p, c->fd0.p,
                                              c->fd1.p, c->fad.p, c->xa32.p, c->da32.p, &exec);
  if (rc) return rc;
  // more iterations of the running inner solves (raw: plain launches, as
  // inside a capture)
  auto more = [&](int nbat, bool raw) -> int {
    return enqueue_batches<float, float, float>(c, raw ? nullptr : exec, nbat, delta, c->fr.p,
                                                c->fd0.p, c->fd1.p, c->fad.p, c->xa32.p,
                                                c->da32.p);
  };
  // outer step: x - delta into the other solution buffer, r = A x - b with
  // its partial sums, the inner right-hand side -r; then the test
  // (the first one of a solve, first = true: r = A x - b with x as given,
  // no delta to apply and no buffer switch)
  auto outer_step = [&](bool first = false) -> int {
    EbeArgs<double, double> a = base_args<double, double>(c, c->xa.p, c->da.p);
    a.b = b;
    a.r32 = c->fr.p;
    a.partial = part_u(c, 0);
    a.skip_if_outer_done = 1;
    if (first) {
      a.y = x;
      HIP_OK((launch_ebe<double, double, STAGE_PLAIN, EPI_RESID_MIXED, FIN_PARTIAL>(
          c, a, false, c->typdia, c->typext)));
    } else {
      a.xb0 = x;
      a.xb1 = c->vx2.p;
      a.delta = delta;
      HIP_OK((launch_ebe<double, double, STAGE_XUPD, EPI_RESID_MIXED, FIN_PARTIAL>(
          c, a, false, c->typdia, c->typext)));
    }
    hipLaunchKernelGGL(k_outer_check, dim3(c->nsys), dim3(kThreads), 0, c->stream,
                       part_u(c, 0), L.nblocks, sys_ptr(c->sys_blocks), c->st_p,
                       first ? 0 : 1);
    HIP_OK(hipGetLastError());
    return BHIP_OK;
  };
  const bool reliable = c->reliable;
  // the first kernel of the inner solves an outer test has just started
  // (reliable updates: a continued solve reads its previous direction from
  // fdz, zero for a fresh start, where BETA = 0)
  auto inner_start = [&]() -> int {
    if (reliable)
      hipLaunchKernelGGL(k_inner_dir, dim3(L.nblocks), dim3(kThreads), 0, c->stream, c->fdz.p,
                         c->fd0.p, c->fd1.p, c->node_start.p, sys_ptr(c->block_sys), c->st_p);
    HIP_OK((enqueue_cg_start<float, float>(c, c->fr.p, c->fd0.p, c->fad.p, c->fdz.p, c->xa32.p,
                                           c->da32.p, L.nblocks, delta)));
    return BHIP_OK;
  };
  int& hint_in = c->iter_hint[1];   // iterations of one inner solve
  int& hint_out = c->iter_hint[2];  // inner solves of one mixed solve
  const int nout = std::min(std::max(hint_out, 1), 8);
  // iterations queued for inner solve k: what inner solve k of the previous
  // solve needed, rounded up to even (later batches keep the parity), or the
  // average inner count when that is unknown
  int npl[8];
  uint64_t plan_key[2] = {0, 0};
  for (int k = 0; k < nout; ++k) {
    const int h = c->inner_hint[k] > 0 ? c->inner_hint[k] : std::max(hint_in, kBatch);
    npl[k] = std::min(even_up(h), 0xfffe);
    plan_key[k / 4] |= (uint64_t)npl[k] << (16 * (k % 4));
  }
  const bool da32_needed = c->typdia == 'Q';
  // the planned part: FP32 copies, set-up, then the outer and inner steps
  // the previous solve needed (the device skips what is not needed)
  auto plan = [&](bool capturing) -> int {
    (void)capturing;
    if (!xa32_ready)
      hipLaunchKernelGGL(k_to_f32, dim3(1024), dim3(kThreads), 0, c->stream, c->xa32.p, c->xa.p,
                         ncf);
    if (da32_needed)
      hipLaunchKernelGGL(k_to_f32, dim3(g), dim3(kThreads), 0, c->stream, c->da32.p, c->da.p,
                         (long)n);
    HIP_OK(hipGetLastError());
    HIP_OK(enqueue_bb_setup(c, b, eps, nitmax, 1, inner_eps, inner_nitmax, reliable ? 1 : 0));
    int r = outer_step(true);
    for (int k = 0; k < nout && !r; ++k) {
      if ((r = inner_start())) break;
      HIP_OK((enqueue_cg_batch<float, float, float>(c, npl[k], delta, c->fr.p, c->fd0.p,
                                                    c->fd1.p, c->fad.p, c->xa32.p,
                                                    c->da32.p)));
      r = outer_step();
    }
    return r;
  };
  const uint64_t key[8] = {
      1u | (xa32_ready ? 4u : 0u) | (reliable ? 8u : 0u) | (da32_needed ? 16u : 0u) |
          ((uint64_t)(uint32_t)inner_nitmax << 8),
      (uint64_t)nout | ((uint64_t)(uint32_t)nitmax << 32),
      matrix_tag(c), dbits(eps), dbits(inner_eps), (uint64_t)(uintptr_t)x,
      plan_key[0], plan_key[1]};
  if ((rc = run_planned(c, key, plan))) return rc;
  // further rounds while a system has not stopped: start the inner solves
  // the last outer tests began, iterate the running ones, test again. Every
  // round runs at least kBatch inner iterations of each running solve, and
  // NITMAX (and kMaxOuter) stop every system, so the loop ends.
  int nround = 1, rounds = 0;
  const int nbat = std::max(1, (hint_in + 1 + kBatch - 1) / kBatch);
  for (;;) {
    if ((rc = tail())) return rc;
    if (read_state(c)) return BHIP_EDEVICE;
    bool all = true;
    for (int sy = 0; sy < c->nsys; ++sy) all = all && sh[sy].outer_done != 0;
    if (all) break;
    if (++rounds > 100000) return BHIP_EDEVICE;
    if ((rc = inner_start())) return rc;
    if ((rc = more(nround, false))) return rc;
    if ((rc = outer_step())) return rc;
    nround = std::min(nround * 2, nbat);
  }
  int maxin = 0, maxout = 0;
  long long sumin = 0, sumout = 0;
  for (int sy = 0; sy < c->nsys; ++sy) {
    maxin = std::max(maxin, sh[sy].inner_total);
    maxout = std::max(maxout, sh[sy].outer_count);
    sumin += sh[sy].inner_total;
    sumout += sh[sy].outer_count;
  }
  if (sumout > 0) hint_in = (int)((sumin + sumout - 1) / sumout);
  hint_out = maxout;
  for (int k = 0; k < 8; ++k) {
    int h = 0;
    for (int sy = 0; sy < c->nsys; ++sy)
      if (!sh[sy].zero_rhs && k < sh[sy].outer_count) h = std::max(h, sh[sy].inner_pairs[k]);
    c->inner_hint[k] = h;
  }
  (void)maxin;
  return collect_infos(c, x, true, info);
}

}  // namespace bhip

extern "C" {

}  // extern "C"

#include "krylov.inl"

extern "C" {

}  // extern "C"

namespace bhip {
// out[i] = in[i] (* s[i]); no alignment assumed beyond that of a double
__global__ void k_copy_scaled(double* __restrict__ out, const double* __restrict__ in,
                              const double* __restrict__ s, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    out[i] = s ? in[i] * s[i] : in[i];
}
}  // namespace bhip

// SOLVE on device vectors, in the caller's node order or (internal) in the
// context's own order, which skips the permutations.
static int solve_dev(bhip_ctx* c, const bhip_solver_cfg* cfg, double* x_dev, double* b_dev,
                     bhip_solve_info* info, bool internal, bool from_saved = false) {
  if (!c || !cfg || !x_dev || !b_dev) return BHIP_EBADARG;
  if (from_saved ? !c->have_saved : !c->have_matrix) return BHIP_EBADARG;
  bhip_solve_info dummy;
  if (!info) info = &dummy;
  std::memset(info, 0, sizeof(*info));
  int precon = cfg->precon == 0 ? 1 : cfg->precon;
  if (precon % 3 == 0 || precon % 7 == 0 || precon % 11 == 0 || precon % 13 == 0 ||
      precon % 17 == 0)
    return BHIP_EBADARG;
  if (cfg->slv != 1 && cfg->slv != 6 && cfg->slv != 7) return BHIP_EBADARG;
  if (cfg->slv == 7 && (cfg->krylov < 1 || cfg->krylov > 20)) return BHIP_EBADARG;
  if (cfg->slv == 6 && cfg->precision == BHIP_PREC_MIXED) return BHIP_EBADARG;
  if (cfg->precision != BHIP_PREC_FP64 && cfg->precision != BHIP_PREC_MIXED) return BHIP_EBADARG;
  // ensembles: CG, FP64 or mixed (per-system scalars in its kernels)
  if (c->nsys > 1 && cfg->slv != 1) return BHIP_EBADARG;
  const Layout& L = c->L;
  const int n = L.npoin;
  const int g = c->grid_nodes;
  const bool diag = (precon % 2 == 0) || (precon % 5 == 0);
  // OM 'M=N' from the saved copy first: with diagonal preconditioning and
  // off-diagonal terms, PRECD1 and D A D read the saved copy and write the
  // context's matrix, which takes the place of the copy; otherwise the copy
  // is made as it is (a TYPEXT '0' matrix is stored as zero coefficients,
  // which D A D does not write)
  const double* da_src = c->da.p;
  const double* xa_src = c->xa.p;
  if (from_saved && diag && c->ncoef_saved > 0) {
    c->typdia = c->typdia_saved;
    c->typext = c->typext_saved;
    c->ncoef = c->ncoef_saved;
    c->have_matrix = true;
    da_src = c->da_saved.p;
    xa_src = c->xa_saved.p;
  } else if (from_saved) {
    const int rr = bhip_matrix_restore(c);
    if (rr) return rr;
  }
  // caller order -> internal order (with diagonal preconditioning the
  // PRECD1 kernel below reads the caller's vectors itself)
  if (!diag && internal) {
    hipLaunchKernelGGL(k_copy_scaled, dim3(g), dim3(kThreads), 0, c->stream, c->vx.p, x_dev,
                       nullptr, n);
    hipLaunchKernelGGL(k_copy_scaled, dim3(g), dim3(kThreads), 0, c->stream, c->vb.p, b_dev,
                       nullptr, n);
  } else if (!diag) {
    hipLaunchKernelGGL(k_gather<double>, dim3(g), dim3(kThreads), 0, c->stream, c->vx.p, x_dev,
                       c->perm.p, n);
    hipLaunchKernelGGL(k_gather<double>, dim3(g), dim3(kThreads), 0, c->stream, c->vb.p, b_dev,
                       c->perm.p, n);
  }
  HIP_OK(hipGetLastError());
  const bool mixed = cfg->precision == BHIP_PREC_MIXED;
  if (mixed && ensure_f32(c)) return BHIP_EDEVICE;
  bool xa32_ready = false;
  if (diag) {
    // PRECD1 (precd1.f:90-149), X and B taken from the caller's vectors
    hipLaunchKernelGGL(k_precd1_nodes, dim3(g), dim3(kThreads), 0, c->stream, c->vx.p, c->vb.p,
                       x_dev, b_dev, internal ? nullptr : c->perm.p, da_src, c->da.p, c->vdinv.p,
                       n, precon, c->typdia == 'Q', c->typdia == 'I');
    HIP_OK(hipGetLastError());
    if (c->typdia == 'I') c->typdia = 'Q';
    if (c->ncoef > 0) {
      // the mixed path takes its FP32 copy of D A D in the same pass
      float* x32 = mixed ? c->xa32.p : nullptr;
      HIP_OK(launch_dmd(c, x32, c->vdinv.p, xa_src));
      xa32_ready = mixed;
    }
    if (precon % 2 == 0 || precon % 3 == 0) c->typdia = 'I';
  }
  int rc;
  // X <- X*D (solve.f:637-641) and B <- D*B (as the reference leaves it),
  // back to the caller's order or copied as they are (internal). A mixed
  // solve of one system reads X from the solution buffer its state names;
  // an ensemble's first gathers each system's into vx.
  const bool mixed_cg = cfg->slv == 1 && mixed;
  auto tail = [&]() -> int {
    const int32_t* idx = internal ? nullptr : c->perm.p;
    const SolveState* stx = nullptr;
    if (mixed_cg && c->nsys == 1) {
      stx = c->st_p;
    } else if (mixed_cg) {
      hipLaunchKernelGGL(k_xpar_gather, dim3(L.nblocks), dim3(kThreads), 0, c->stream, c->vx.p,
                         c->vx2.p, c->node_start.p, sys_ptr(c->block_sys), c->st_p);
    }
    hipLaunchKernelGGL(k_x_out, dim3(g), dim3(kThreads), 0, c->stream, x_dev, c->vx.p,
                       c->vx2.p, stx, diag ? c->vdinv.p : nullptr, idx, n);
    hipLaunchKernelGGL(k_x_out, dim3(g), dim3(kThreads), 0, c->stream, b_dev, c->vb.p,
                       nullptr, nullptr, nullptr, idx, n);
    HIP_OK(hipGetLastError());
    return BHIP_OK;
  };
  if (cfg->slv == 6) {
    rc = bhip::cgstab_fp64(c, c->vx.p, c->vb.p, cfg->eps, cfg->nitmax, info);
  } else if (cfg->slv == 7 && mixed) {
    rc = bhip::gmres_mixed(c, c->vx.p, c->vb.p, cfg->eps, cfg->nitmax, cfg->krylov, xa32_ready,
                           info);
  } else if (cfg->slv == 7) {
    rc = bhip::gmres_fp64(c, c->vx.p, c->vb.p, cfg->eps, cfg->nitmax, cfg->krylov, info);
  } else if (cfg->precision == BHIP_PREC_MIXED) {
    double ie = cfg->inner_eps > 0 ? cfg->inner_eps : 1e-4;
    int inn = cfg->inner_nitmax > 0 ? cfg->inner_nitmax : std::max(cfg->nitmax, 1);
    rc = cg_mixed(c, c->vx.p, c->vb.p, cfg->eps, std::max(cfg->nitmax, 1), ie, inn, xa32_ready,
                  info, tail);
  } else {
    rc = cg_fp64(c, c->vx.p, c->vb.p, cfg->eps, cfg->nitmax, info, tail);
  }
  if (rc != BHIP_OK && rc != BHIP_MAXITER && rc != BHIP_ZERO_RHS) return rc;
  // the CG drivers wrote the results before their last wait, unless a
  // system with a zero right-hand side has since had its X set to 0
  const bool cg = cfg->slv == 1;
  bool zero_any = false;
  for (int sy = 0; cg && sy < c->nsys; ++sy) zero_any = zero_any || c->st_host[sy].zero_rhs;
  if (!cg || zero_any) {
    const int rt = tail();
    if (rt) return rt;
  }
  return rc;
}

extern "C" {

int bhip_solve_dev(bhip_ctx* c, const bhip_solver_cfg* cfg, double* x_dev, double* b_dev,
                   bhip_solve_info* info) {
  return solve_dev(c, cfg, x_dev, b_dev, info, false);
}

int bhip_solve_dev_internal(bhip_ctx* c, const bhip_solver_cfg* cfg, double* x_dev,
                            double* b_dev, bhip_solve_info* info) {
  return solve_dev(c, cfg, x_dev, b_dev, info, true);
}

int bhip_solve_dev_restored(bhip_ctx* c, const bhip_solver_cfg* cfg, double* x_dev,
                            double* b_dev, bhip_solve_info* info) {
  return solve_dev(c, cfg, x_dev, b_dev, info, false, true);
}

int bhip_solve_dev_noalloc(bhip_ctx* c, const bhip_solver_cfg* cfg, double* x_dev,
                           double* b_dev, bhip_solve_info* info) {
  if (!c || !cfg) return BHIP_EBADARG;
  if (cfg->slv != 1 || cfg->precision != BHIP_PREC_FP64) {
    fprintf(stderr, "libbiefhip: bhip_solve_dev_noalloc runs FP64 CG (slv 1) only\n");
    return BHIP_EBADARG;
  }
  if (c->max_part > kScratchParts || c->nsys > kScratchSys) {
    fprintf(stderr,
            "libbiefhip: bhip_solve_dev_noalloc: %d partial sums and %d systems exceed the "
            "static scratch (%d and %d)\n",
            c->max_part, c->nsys, kScratchParts, kScratchSys);
    return BHIP_EBADARG;
  }
  unsigned char* base = static_cast<unsigned char*>(scratch_addr(c->device));
  if (!base) return BHIP_EDEVICE;
  // this call's partial sums and states in the scratch, graphs off (a graph
  // holds the addresses it was captured with)
  struct Restore {
    bhip_ctx* c;
    double* part;
    SolveState* st;
    bool graphs;
    ~Restore() {
      c->part_p = part;
      c->st_p = st;
      c->use_graphs = graphs;
    }
  } restore{c, c->part_p, c->st_p, c->use_graphs};
  c->part_p = reinterpret_cast<double*>(base);
  c->st_p = reinterpret_cast<SolveState*>(base + kScratchPartBytes);
  c->use_graphs = false;
  // the scratch keeps the previous call's contents: clear what this call uses
  const long nd = (long)4 * c->max_part;
  const long ns = (long)(sizeof(SolveState) * c->nsys / sizeof(double));
  hipLaunchKernelGGL(k_fill<double>, dim3(grid_for((int)nd)), dim3(kThreads), 0, c->stream,
                     c->part_p, 0.0, nd);
  hipLaunchKernelGGL(k_fill<double>, dim3(1), dim3(kThreads), 0, c->stream,
                     reinterpret_cast<double*>(c->st_p), 0.0, ns);
  HIP_OK(hipGetLastError());
  return solve_dev(c, cfg, x_dev, b_dev, info, false);
}

int bhip_ctx_perm(const bhip_ctx* c, int32_t* perm) {
  if (!c || !perm) return BHIP_EBADARG;
  std::memcpy(perm, c->L.perm.data(), sizeof(int32_t) * c->L.perm.size());
  return BHIP_OK;
}

}  // extern "C"

namespace bhip {
// Reads every line of a buffer larger than the last-level cache, so that the
// kernel timed next finds none of its data cached (and no dirty lines of the
// flush to write back); the sum is stored only if it is NaN, which keeps the
// reads without writing anything for a finite buffer.
__global__ void k_flush(double* __restrict__ x, long n) {
  double s = 0.0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    s += x[i];
  if (s != s) x[0] = s;
}
}  // namespace bhip

// The two kernels of one CG iteration, `reps` launches each, timed with HIP
// events on the context stream: back to back (flush == nullptr), or each
// launch after a pass over `flush` (cold caches; only the kernel is inside
// the events).
static int time_cg(bhip_ctx* c, int precision, int reps, double* flush, long flush_n,
                   double* ms_dir, double* ms_upd) {
  if (!c || !c->have_matrix || reps < 1) return BHIP_EBADARG;
  if (precision == BHIP_PREC_MIXED && !c->fr.p) return BHIP_EBADARG;
  if (flush && flush_n < 1) return BHIP_EBADARG;
  // never converges, never reaches NITMAX: every launch does its full work
  for (int sy = 0; sy < c->nsys; ++sy) {
    SolveState* sh = c->st_host + sy;
    std::memset(sh, 0, sizeof(SolveState));
    init_cg_state(sh, -1.0, 0, 0x7fffffff);
    sh->m[0] = sh->m[1] = 1;
  }
  if (write_state(c)) return BHIP_EDEVICE;
  hipEvent_t e0, e1;
  HIP_OK(hipEventCreate(&e0));
  HIP_OK(hipEventCreate(&e1));
  float t_dir = 0.f, t_upd = 0.f;
  auto time_it = [&](auto&& launch, float* out) -> int {
    if (!flush) {
      HIP_OK(hipEventRecord(e0, c->stream));
      for (int i = 0; i < reps; ++i) HIP_OK(launch());
      HIP_OK(hipEventRecord(e1, c->stream));
      HIP_OK(hipEventSynchronize(e1));
      HIP_OK(hipEventElapsedTime(out, e0, e1));
      return BHIP_OK;
    }
    *out = 0.f;
    for (int i = 0; i < reps; ++i) {
      hipLaunchKernelGGL(k_flush, dim3(4096), dim3(kThreads), 0, c->stream, flush, flush_n);
      HIP_OK(hipGetLastError());
      HIP_OK(hipEventRecord(e0, c->stream));
      HIP_OK(launch());
      HIP_OK(hipEventRecord(e1, c->stream));
      HIP_OK(hipEventSynchronize(e1));
      float t = 0.f;
      HIP_OK(hipEventElapsedTime(&t, e0, e1));
      *out += t;
    }
    return BHIP_OK;
  };
  int rc;
  if (precision == BHIP_PREC_MIXED) {
    EbeArgs<float, float> a = base_args<float, float>(c, c->xa32.p, c->da32.p);
    a.r = c->fr.p;
    a.dold = c->fd0.p;
    a.dnew = c->fd1.p;
    a.out = c->fad.p;
    a.partial = part_e(c, 0);
    a.part_in = part_u(c, 0);
    a.n_part_in = c->L.nblocks;
    a.parity = 0;
    rc = time_it([&]() {
      return launch_ebe<float, float, STAGE_CGDIR, EPI_CGAD, FIN_CG>(c, a, false, c->typdia,
                                                                      c->typext);
    }, &t_dir);
    if (rc) return rc;
    if (ms_upd) {
      rc = time_it([&]() {
        return launch_update<float, float>(c, c->fx0.p, c->fr.p, c->fd1.p, c->fad.p, 1);
      }, &t_upd);
      if (rc) return rc;
    }
  } else {
    EbeArgs<double, double> a = base_args<double, double>(c, c->xa.p, c->da.p);
    a.r = c->vr.p;
    a.dold = c->vd0.p;
    a.dnew = c->vd1.p;
    a.out = c->vad.p;
    a.partial = part_e(c, 0);
    a.part_in = part_u(c, 0);
    a.n_part_in = c->L.nblocks;
    a.parity = 0;
    rc = time_it([&]() {
      return launch_ebe<double, double, STAGE_CGDIR, EPI_CGAD, FIN_CG>(c, a, false, c->typdia,
                                                                        c->typext);
    }, &t_dir);
    if (rc) return rc;
    if (ms_upd) {
      rc = time_it([&]() {
        return launch_update<double, double>(c, c->vio0.p, c->vr.p, c->vd1.p, c->vad.p, 1);
      }, &t_upd);
      if (rc) return rc;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (ms_dir) *ms_dir = t_dir / reps;
  if (ms_upd) *ms_upd = t_upd / reps;
  return BHIP_OK;
}

extern "C" {

int bhip_time_cg_kernels(bhip_ctx* c, int precision, int reps, double* ms_dir,
                         double* ms_upd) {
  double upd = 0.0;
  const int rc = time_cg(c, precision, reps, nullptr, 0, ms_dir, &upd);
  if (ms_upd) *ms_upd = upd;
  return rc;
}

int bhip_time_cg_dir_cold(bhip_ctx* c, int precision, int reps, double* flush_dev,
                          long flush_n, double* ms_dir) {
  if (!flush_dev || !ms_dir) return BHIP_EBADARG;
  return time_cg(c, precision, reps, flush_dev, flush_n, ms_dir, nullptr);
}

int bhip_time_matvec(bhip_ctx* c, int precision, int reps, double* ms) {
  if (!c || !c->have_matrix || reps < 1 || !ms) return BHIP_EBADARG;
  if (precision == BHIP_PREC_MIXED && !c->fr.p) return BHIP_EBADARG;
  hipEvent_t e0, e1;
  HIP_OK(hipEventCreate(&e0));
  HIP_OK(hipEventCreate(&e1));
  HIP_OK(hipEventRecord(e0, c->stream));
  for (int i = 0; i < reps; ++i) {
    if (precision == BHIP_PREC_MIXED) {
      EbeArgs<float, float> a = base_args<float, float>(c, c->xa32.p, c->da32.p);
      a.y = c->fr.p;
      a.out = c->fad.p;
      a.op = BHIP_OP_AY;
      HIP_OK((launch_ebe<float, float, STAGE_PLAIN, EPI_OP, FIN_NONE>(c, a, false, c->typdia,
                                                                      c->typext)));
    } else {
      EbeArgs<double, double> a = base_args<double, double>(c, c->xa.p, c->da.p);
      a.y = c->vio0.p;
      a.out = c->vio1.p;
      a.op = BHIP_OP_AY;
      HIP_OK((launch_ebe<double, double, STAGE_PLAIN, EPI_OP, FIN_NONE>(c, a, false, c->typdia,
                                                                        c->typext)));
    }
  }
  HIP_OK(hipEventRecord(e1, c->stream));
  HIP_OK(hipEventSynchronize(e1));
  float t = 0.f;
  HIP_OK(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms = t / reps;
  return BHIP_OK;
}

int bhip_solve(bhip_ctx* c, const bhip_solver_cfg* cfg, double* x, double* b,
               bhip_solve_info* info) {
  if (!c || !x || !b) return BHIP_EBADARG;
  const int n = c->L.npoin;
  const size_t bytes = sizeof(double) * n;
  HIP_OK(c->h_x.alloc(n));
  HIP_OK(c->h_b.alloc(n));
  HIP_OK(to_device(c, c->h_x.p, x, bytes));
  HIP_OK(to_device(c, c->h_b.p, b, bytes));
  int rc = bhip_solve_dev(c, cfg, c->h_x.p, c->h_b.p, info);
  if (rc == BHIP_OK || rc == BHIP_MAXITER || rc == BHIP_ZERO_RHS) {
    HIP_OK(to_host(c, x, c->h_x.p, bytes));
    HIP_OK(to_host(c, b, c->h_b.p, bytes));
  }
  (void)hipStreamSynchronize(c->stream);
  return rc;
}

}  // extern "C"

// ===========================================================================
// operand preparation: DIRICH, LUMP, element matrices
// ===========================================================================
namespace bhip {

// DIRAUX (diraux.f) after SM -= A FBOR (diri01.f:94-116): at prescribed
// nodes SM = DA*FBOR, W = 0, F = FBOR; elsewhere W = 1.
__global__ void k_diraux(double* __restrict__ sm, const double* __restrict__ ax,
                         const double* __restrict__ da, const double* __restrict__ fbor,
                         const int32_t* __restrict__ limdir, int kdir, double* __restrict__ w,
                         double* __restrict__ f, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double s = sm[i] - ax[i];
    double wi = 1.0;
    if (limdir[i] == kdir) {
      s = da[i] * fbor[i];
      wi = 0.0;
      f[i] = fbor[i];
    }
    sm[i] = s;
    w[i] = wi;
  }
}

template <typename T>
__global__ void k_gather_any(T* __restrict__ out, const T* __restrict__ in,
                             const int32_t* __restrict__ idx, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    out[i] = in[idx[i]];
}

// Element geometry of slot s: XEL, YEL of nodes 2 and 3 (node 1 at the
// origin, inbief.f:384-397) and SURFAC. From the context's slot-ordered copy
// (geo: [5][nslot_pad], one coalesced read per value) when it is given, else
// from the caller's arrays by element e (a gather per value).
struct SlotGeo {
  double x2, x3, y2, y3, s;
};
__device__ inline SlotGeo load_geo(const double* __restrict__ geo, size_t np, int s,
                                   const double* __restrict__ xel,
                                   const double* __restrict__ yel,
                                   const double* __restrict__ surfac, size_t ne, int e,
                                   bool xy) {
  SlotGeo g = {0.0, 0.0, 0.0, 0.0, 0.0};
  if (geo) {
    g.s = geo[4 * np + s];
    if (xy) {
      g.x2 = geo[s];
      g.x3 = geo[np + s];
      g.y2 = geo[2 * np + s];
      g.y3 = geo[3 * np + s];
    }
  } else {
    g.s = surfac[e];
    if (xy) {
      g.x2 = xel[ne + e];
      g.x3 = xel[2 * ne + e];
      g.y2 = yel[ne + e];
      g.y3 = yel[2 * ne + e];
    }
  }
  return g;
}

// geo[c][s] for c = XEL2, XEL3, YEL2, YEL3, SURFAC of slot s's element
__global__ void k_geo_slots(double* __restrict__ geo, const double* __restrict__ xel,
                            const double* __restrict__ yel, const double* __restrict__ surfac,
                            const int32_t* __restrict__ slot_elem, int nslot, int nslot_pad,
                            int nelem) {
  const size_t np = nslot_pad, ne = nelem;
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < nslot; s += gridDim.x * blockDim.x) {
    const int e = slot_elem[s];
    geo[s] = xel[ne + e];
    geo[np + s] = xel[2 * ne + e];
    geo[2 * np + s] = yel[ne + e];
    geo[3 * np + s] = yel[2 * ne + e];
    geo[4 * np + s] = surfac[e];
  }
}

// Element matrices per slot (each element instance computes the same values)
// and the assembly of their diagonal terms onto owned nodes in LDS, colour
// by colour (same deterministic scheme as the mat-vec).
//   kind 100: MT01AA mass (mt01aa.f:83-106)
//   kind 1  : MT02AA P1 isotropic viscosity (mt02aa.f:247-290)
//   kind 2  : MT02AA 'UV' (mt02aa.f:200-240)
// accumulate: OM 'M=M+N' (else 'M=N').
__global__ void __launch_bounds__(kThreads)
k_elem_tri(int kind, double xmul, const double* __restrict__ surfac,
           const double* __restrict__ xel, const double* __restrict__ yel, int nelem,
           const double* __restrict__ u_int, const double* __restrict__ v_int,
           const int32_t* __restrict__ slot_elem, const uint16_t* __restrict__ lid,
           int nslot_pad, const int32_t* __restrict__ node_start,
           const int32_t* __restrict__ halo_start, const int32_t* __restrict__ halo,
           const int32_t* __restrict__ color_off, int nblocks, double* __restrict__ xa,
           double* __restrict__ da, int accumulate, int target_q,
           const double* __restrict__ geo) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double* acc = reinterpret_cast<double*>(smem_raw);
  const int b = blockIdx.x;
  if (b >= nblocks) return;
  const int n0 = node_start[b], nown = node_start[b + 1] - n0;
  const int h0 = halo_start[b];
  for (int l = threadIdx.x; l < nown; l += blockDim.x) acc[l] = 0.0;
  __syncthreads();
  const int32_t* co = color_off + (size_t)b * (kMaxColors + 1);
  const int send = co[kMaxColors];
  for (int c = 0; c < kMaxColors; ++c) {
    const int s0 = co[c];
    if (s0 >= send) break;
    const int s1 = co[c + 1];
    for (int s = s0 + threadIdx.x; s < s1; s += blockDim.x) {
      const int e = slot_elem[s];
      int l[3], g[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        l[k] = lid[(size_t)k * nslot_pad + s];
        g[k] = l[k] < nown ? n0 + l[k] : halo[h0 + l[k] - nown];
      }
      const SlotGeo G = load_geo(geo, nslot_pad, s, xel, yel, surfac, nelem, e, kind != 100);
      double a12, a13, a23, t1, t2, t3;
      if (kind == 100) {
        const double det = G.s * (xmul / 12.0);
        a12 = a13 = a23 = det;
        t1 = t2 = t3 = det + det;
      } else {
        const double x2 = G.x2, x3 = G.x3;
        const double y2 = G.y2, y3 = G.y3;
        if (kind == 2) {
          const double g1 = v_int[g[0]], g2 = v_int[g[1]], g3 = v_int[g[2]];
          const double coef1 = (xmul / 48.0) / G.s;
          const double g123 = g1 + g2 + g3;
          const double coef2 = u_int[g[0]] * (g1 + g123) + u_int[g[1]] * (g2 + g123) +
                               u_int[g[2]] * (g3 + g123);
          a12 = (y3 * (y2 - y3) + x3 * (x2 - x3)) * coef2 * coef1;
          a13 = -(y2 * (y2 - y3) + x2 * (x2 - x3)) * coef2 * coef1;
          a23 = -(y2 * y3 + x2 * x3) * coef2 * coef1;
        } else {
          const double somvx = (u_int[g[0]] + u_int[g[1]] + u_int[g[2]]) * (xmul / 12.0) /
                               G.s;
          const double x2x3 = x2 * x3 + y2 * y3;
          const double x2aux = x2 * (-x2 + x3) + y2 * (-y2 + y3);
          const double x3aux = x3 * (-x2 + x3) + y3 * (-y2 + y3);
          a12 = -somvx * x3aux;
          a13 = somvx * x2aux;
          a23 = -somvx * x2x3;
        }
        t1 = -a12 - a13;
        t2 = -a12 - a23;
        t3 = -a13 - a23;
      }
      // a symmetric element matrix; into a non-symmetric ('Q') target it
      // goes to both halves (OM 'M=M+N' of an 'S' N into a 'Q' M)
      const double av[3] = {a12, a13, a23};
      for (int cc = 0; cc < (target_q ? 6 : 3); ++cc) {
        double* p = xa + (size_t)cc * nslot_pad + s;
        *p = accumulate ? *p + av[cc % 3] : av[cc % 3];
      }
      const double tk[3] = {t1, t2, t3};
#pragma unroll
      for (int k = 0; k < 3; ++k)
        if (l[k] < nown) acc[l[k]] += tk[k];
    }
    __syncthreads();
  }
  for (int l = threadIdx.x; l < nown; l += blockDim.x) {
    if (accumulate)
      da[n0 + l] += acc[l];
    else
      da[n0 + l] = acc[l];
  }
}

// One horizontal direction of the prism diffusion matrix (mt02pp.f:229-356
// along x with F, 358-445 along y with G): d1, d2, d3 the in-plane factors of
// the base triangle, n[6] the nodal coefficients over the column heights,
// hh the heights, zz the node elevations, dd = XMUL / 2880 / SURFAC. Adds
// (or stores) the 15 pair contributions into c.
__device__ inline void mt02pp_dir(double d1, double d2, double d3, const double (&n)[6],
                                  const double (&hh)[3], const double (&zz)[6], double dd,
                                  double (&c)[15], bool add) {
  const double h1 = hh[0], h2 = hh[1], h3 = hh[2];
  const double sh1 = h1 + h2 + h3;
  const double shh = h1 * h1 + h2 * h2 + h3 * h3;
  const double ri = -(zz[0] * d1 + zz[1] * d2 + zz[2] * d3);
  const double rs = -(zz[3] * d1 + zz[4] * d2 + zz[5] * d3);
  const double r = ri + rs;
  const double rri = r + ri + ri, rrs = r + rs + rs;
  const double rrri = r * r + 2 * ri * ri, rrrs = r * r + 2 * rs * rs;
  const double d12 = d1 * d2, d13 = d1 * d3, d23 = d2 * d3;
  const double sni = n[0] + n[1] + n[2], sns = n[3] + n[4] + n[5];
  const double snhi = n[0] * h1 + n[1] * h2 + n[2] * h3;
  const double snhs = n[3] * h1 + n[4] * h2 + n[5] * h3;
  const double shi = (sni * sh1 + snhi + snhi) * sh1 + sni * shh +
                     2 * (n[0] * h1 * h1 + n[1] * h2 * h2 + n[2] * h3 * h3);
  const double shs = (sns * sh1 + snhs + snhs) * sh1 + sns * shh +
                     2 * (n[3] * h1 * h1 + n[4] * h2 * h2 + n[5] * h3 * h3);
  const double snhh = shi + shs;
  const double snhhi = snhh + shi + shi, snhhs = snhh + shs + shs;
  const double snhi1 = (sni + n[0]) * (sh1 + h1) + snhi + n[0] * h1;
  const double snhs1 = (sns + n[3]) * (sh1 + h1) + snhs + n[3] * h1;
  const double snhi2 = (sni + n[1]) * (sh1 + h2) + snhi + n[1] * h2;
  const double snhs2 = (sns + n[4]) * (sh1 + h2) + snhs + n[4] * h2;
  const double snhi3 = (sni + n[2]) * (sh1 + h3) + snhi + n[2] * h3;
  const double snhs3 = (sns + n[5]) * (sh1 + h3) + snhs + n[5] * h3;
  const double hri1 = rri * snhi1 + r * snhs1, hrs1 = rrs * snhs1 + r * snhi1;
  const double hri2 = rri * snhi2 + r * snhs2, hrs2 = rrs * snhs2 + r * snhi2;
  const double hri3 = rri * snhi3 + r * snhs3, hrs3 = rrs * snhs3 + r * snhi3;
  const double rr11 = 2 * (rrri * (sni + n[0] + n[0]) + rrrs * (sns + n[3] + n[3]));
  const double rr22 = 2 * (rrri * (sni + n[1] + n[1]) + rrrs * (sns + n[4] + n[4]));
  const double rr33 = 2 * (rrri * (sni + n[2] + n[2]) + rrrs * (sns + n[5] + n[5]));
  const double rr12 = rrri * (sni + n[0] + n[1]) + rrrs * (sns + n[3] + n[4]);
  const double rr13 = rrri * (sni + n[0] + n[2]) + rrrs * (sns + n[3] + n[5]);
  const double rr23 = rrri * (sni + n[1] + n[2]) + rrrs * (sns + n[4] + n[5]);
  const double hh12 = snhh * d12, hh13 = snhh * d13, hh23 = snhh * d23;
  double v[15];
  v[0] = dd * (snhhi * d12 - d1 * hri2 - d2 * hri1 + rr12);
  v[1] = dd * (snhhi * d13 - d1 * hri3 - d3 * hri1 + rr13);
  v[2] = dd * (-hh12 - hh13 + d1 * (hri1 - hrs1) - rr11);
  v[3] = dd * (hh12 + d1 * hri2 - d2 * hrs1 - rr12);
  v[4] = dd * (hh13 + d1 * hri3 - d3 * hrs1 - rr13);
  v[5] = dd * (snhhi * d23 - d2 * hri3 - d3 * hri2 + rr23);
  v[6] = dd * (hh12 + d2 * hri1 - d1 * hrs2 - rr12);
  v[7] = dd * (-hh12 - hh23 + d2 * (hri2 - hrs2) - rr22);
  v[8] = dd * (hh23 + d2 * hri3 - d3 * hrs2 - rr23);
  v[9] = dd * (hh13 + d3 * hri1 - d1 * hrs3 - rr13);
  v[10] = dd * (hh23 + d3 * hri2 - d2 * hrs3 - rr23);
  v[11] = dd * (-hh13 - hh23 + d3 * (hri3 - hrs3) - rr33);
  v[12] = dd * (snhhs * d12 + d1 * hrs2 + d2 * hrs1 + rr12);
  v[13] = dd * (snhhs * d13 + d1 * hrs3 + d3 * hrs1 + rr13);
  v[14] = dd * (snhhs * d23 + d2 * hrs3 + d3 * hrs2 + rr23);
#pragma unroll
  for (int q = 0; q < 15; ++q) c[q] = add ? c[q] + v[q] : v[q];
}

// the pairs (of the 15, row-major upper order) holding each local node
__constant__ int kPrismRowPairs[6][5] = {{0, 1, 2, 3, 4},   {0, 5, 6, 7, 8},
                                        {1, 5, 9, 10, 11},  {2, 6, 9, 12, 13},
                                        {3, 7, 10, 12, 14}, {4, 8, 11, 13, 14}};

// inputs of one prism slot: ids, geometry, current coefficients for M=M+N
struct PrismIn {
  int s;           // slot; -1 when the thread has none in the phase
  int l[6];
  SlotGeo G;
  double old[15];  // target coefficients before 'M=M+N' (the first 15)
};

// Prism element matrices per slot, diagonal assembled in LDS colour by colour.
//   KIND 200: MT01PP mass (mt01pp.f: XMUL*SURFAC/360 combinations of the
//             column heights floored at 1e-3); nd holds Z per node
//   KIND 201: MT02PP diffusion (mt02pp.f), P1 coefficients F, G, H; nd holds
//             (Z, F, G, H + H_up) per node, H_up being H at the node NPOU0
//             above (the caller's II4 = I1 + NPOU0); h_zero cancels the
//             vertical term, inchyd the hydrostatic inconsistency test, mon
//             the 'MON' clipping
// ST 2 copies the records of the block's local nodes (owned, then halo) into
// LDS once, after the diagonal accumulator (rec_off doubles in), and the
// slots read them there; ST 0 gathers each slot's six records from global
// memory. Both give the same bits. A workgroup of the staged MT02PP holds up
// to 160 KB of LDS, so it may be alone on its CU: the staged form therefore
// runs 2 * kThreads threads in two groups, each computing the slots of its
// own colour, and adds their diagonal terms in colour order. ACC 0 writes the
// element matrices, 1 adds them to a symmetric target ('M=M+N'), 2 to both
// halves of a non-symmetric one.
template <int KIND, int ST, int ACC>
__global__ void __launch_bounds__(ST == 2 ? 2 * kThreads : kThreads)
k_elem_prism(double xmul, const double* __restrict__ surfac,
             const double* __restrict__ xel, const double* __restrict__ yel, int nelem,
             const double* __restrict__ nd, int h_zero, int inchyd, int mon,
             const int32_t* __restrict__ slot_elem, const uint16_t* __restrict__ lid,
             int nslot_pad, const int32_t* __restrict__ node_start,
             const int32_t* __restrict__ halo_start, const int32_t* __restrict__ halo,
             const int32_t* __restrict__ color_off, int nblocks, double* __restrict__ xa,
             double* __restrict__ da, const double* __restrict__ geo, int rec_off,
             int col_ids) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double* acc = reinterpret_cast<double*>(smem_raw);
  double* rec = acc + rec_off;
  constexpr int W = KIND == 200 ? 1 : 4;  // doubles per node record
  const int b = blockIdx.x;
  if (b >= nblocks) return;
  const int n0 = node_start[b], nown = node_start[b + 1] - n0;
  const int h0 = halo_start[b];
  const int32_t* co = color_off + (size_t)b * (kMaxColors + 1);
  const int send = co[kMaxColors];
  const size_t np = nslot_pad;
  // ST 2: two groups of kThreads threads, each on its own colour
  constexpr int NT = ST == 2 ? 2 * kThreads : kThreads;
  const int tg = (int)threadIdx.x % kThreads;
  const int grp = (int)threadIdx.x / kThreads;

  auto load_slot = [&](PrismIn& in, int s) {
    in.s = s;
    if (s < 0) return;
    // MT01PP reads all six ids: its three-id form measured slower (445 against
    // 382 us on 10.5 M slots), MT02PP's the same either way
    load_ids<6>(in.l, lid, np, s, KIND == 201 ? col_ids : 0);
    const int e = geo ? 0 : slot_elem[s];
    in.G = load_geo(geo, np, s, xel, yel, surfac, nelem, e, KIND != 200);
    if constexpr (ACC > 0) {
#pragma unroll
      for (int q = 0; q < 15; ++q) in.old[q] = xa[(size_t)q * np + s];
    }
  };
  // the thread's first slot of colour c, or -1
  auto first_slot = [&](int c) {
    if (c >= kMaxColors || co[c] >= send) return -1;
    const int s = co[c] + tg;
    return s < co[c + 1] ? s : -1;
  };

  // the element matrix of a slot into XA, its diagonal terms into t
  auto compute_slot = [&](const PrismIn& in, double (&t)[6]) {
    const int s = in.s;
    const SlotGeo& G = in.G;
    int g[6];
    if constexpr (!ST) {
#pragma unroll
      for (int k = 0; k < 6; ++k)
        g[k] = in.l[k] < nown ? n0 + in.l[k] : halo[h0 + in.l[k] - nown];
    }
    double cf[15];
    if constexpr (KIND == 200) {
      double zz[6];
#pragma unroll
      for (int k = 0; k < 6; ++k) zz[k] = ST ? rec[in.l[k]] : nd[g[k]];
      const double coef = G.s * (xmul / 360.0);
      const double h1 = fmax(zz[3] - zz[0], 1e-3) * coef;
      const double h2 = fmax(zz[4] - zz[1], 1e-3) * coef;
      const double h3 = fmax(zz[5] - zz[2], 1e-3) * coef;
      const double ht = h1 + h2 + h3;
      cf[3] = cf[6] = h1 + h2 + ht;
      cf[4] = cf[9] = h1 + h3 + ht;
      cf[8] = cf[10] = h2 + h3 + ht;
      cf[2] = 4 * h1 + ht + ht;
      cf[7] = 4 * h2 + ht + ht;
      cf[11] = 4 * h3 + ht + ht;
      cf[0] = cf[12] = cf[3] + cf[3];
      cf[1] = cf[13] = cf[4] + cf[4];
      cf[5] = cf[14] = cf[8] + cf[8];
      t[0] = t[3] = cf[2] + cf[2];
      t[1] = t[4] = cf[7] + cf[7];
      t[2] = t[5] = cf[11] + cf[11];
    } else {
      // (Z, F, G, H + H_up) of the six nodes, two 16-byte loads each
      double zz[6], fv[6], gv[6], hs[3];
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        double2 a0, a1;
        if constexpr (ST) {
          a0 = *reinterpret_cast<const double2*>(rec + 4 * in.l[k]);
          a1 = *reinterpret_cast<const double2*>(rec + 4 * in.l[k] + 2);
        } else {
          a0 = *reinterpret_cast<const double2*>(nd + 4 * (size_t)g[k]);
          a1 = *reinterpret_cast<const double2*>(nd + 4 * (size_t)g[k] + 2);
        }
        zz[k] = a0.x;
        fv[k] = a0.y;
        gv[k] = a1.x;
        if (k < 3) hs[k] = a1.y;
      }
      const double hh[3] = {zz[3] - zz[0], zz[4] - zz[1], zz[5] - zz[2]};
      bool dry = hh[0] < 1e-4 || hh[1] < 1e-4 || hh[2] < 1e-4;
      if (inchyd)
        dry = dry || fmax(fmax(zz[0], zz[1]), zz[2]) > fmin(fmin(zz[3], zz[4]), zz[5]);
      double nf[6], ng[6], nh[3];
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        nf[k] = dry ? 0.0 : fv[k] / hh[k % 3];
        ng[k] = dry ? 0.0 : gv[k] / hh[k % 3];
      }
#pragma unroll
      for (int k = 0; k < 3; ++k) nh[k] = dry ? 0.0 : hs[k] / hh[k];
      const double dd = (xmul / 2880.0) / G.s;
      const double x2 = G.x2, x3 = G.x3;
      const double y2 = G.y2, y3 = G.y3;
      mt02pp_dir(y2 - y3, y3, -y2, nf, hh, zz, dd, cf, false);
      mt02pp_dir(x3 - x2, -x3, x2, ng, hh, zz, dd, cf, true);
      const double dz = h_zero ? 0.0 : G.s * (xmul / 6.0);
      cf[2] = cf[2] - dz * nh[0];
      cf[7] = cf[7] - dz * nh[1];
      cf[11] = cf[11] - dz * nh[2];
      if (mon) {
#pragma unroll
        for (int q = 0; q < 15; ++q) cf[q] = xmul > 0.0 ? fmin(cf[q], 0.0) : fmax(cf[q], 0.0);
      }
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        double sum = -cf[kPrismRowPairs[k][0]];
#pragma unroll
        for (int j = 1; j < 5; ++j) sum = sum - cf[kPrismRowPairs[k][j]];
        t[k] = sum;
      }
    }
    if constexpr (ACC > 0) {
#pragma unroll
      for (int q = 0; q < 15; ++q) xa[(size_t)q * np + s] = in.old[q] + cf[q];
      // a symmetric element matrix into a non-symmetric ('Q') target goes to
      // both halves (OM 'M=M+N' of an 'S' N into a 'Q' M)
      if constexpr (ACC == 2) {
#pragma unroll
        for (int q = 0; q < 15; ++q) {
          double* p = xa + (size_t)(15 + q) * np + s;
          *p = *p + cf[q];
        }
      }
    } else {
#pragma unroll
      for (int q = 0; q < 15; ++q) xa[(size_t)q * np + s] = cf[q];
    }
  };
  auto add_slot = [&](const PrismIn& in, const double (&t)[6]) {
#pragma unroll
    for (int k = 0; k < 6; ++k)
      if (in.l[k] < nown) acc[in.l[k]] += t[k];
  };
  auto do_slot = [&](const PrismIn& in) {
    double t[6];
    compute_slot(in, t);
    add_slot(in, t);
  };

  for (int l = threadIdx.x; l < nown; l += blockDim.x) acc[l] = 0.0;
  if constexpr (ST) {
    // records of the local nodes, kStage per thread in flight at once
    constexpr int kStage = 4;
    const int nloc = nown + halo_start[b + 1] - h0;
    for (int l0 = threadIdx.x; l0 < nloc; l0 += kStage * NT) {
      double2 va[kStage], vb[kStage];
#pragma unroll
      for (int u = 0; u < kStage; ++u) {
        const int l = l0 + u * NT;
        const int lc = l < nloc ? l : nloc - 1;
        const size_t gl = lc < nown ? n0 + lc : halo[h0 + lc - nown];
        if constexpr (W == 1) {
          va[u].x = nd[gl];
        } else {
          const double2* src = reinterpret_cast<const double2*>(nd + 4 * gl);
          va[u] = src[0];
          vb[u] = src[1];
        }
      }
#pragma unroll
      for (int u = 0; u < kStage; ++u) {
        const int l = l0 + u * NT;
        if (l < nloc) {
          if constexpr (W == 1) {
            rec[l] = va[u].x;
          } else {
            double2* dst = reinterpret_cast<double2*>(rec + 4 * (size_t)l);
            dst[0] = va[u];
            dst[1] = vb[u];
          }
        }
      }
    }
  }
  __syncthreads();
  if constexpr (ST == 2) {
    // colours c0 and c0 + 1 computed at once by the two groups; their
    // diagonal terms go into the accumulator in colour order, one group
    // after the other, so the sums are those of the one-group form
    for (int c0 = 0; c0 < kMaxColors; c0 += 2) {
      if (co[c0] >= send) break;
      const int c = c0 + grp;
      const bool has = c < kMaxColors && co[c] < send;
      PrismIn cur;
      double t[6];
      load_slot(cur, first_slot(c));
      if (cur.s >= 0) compute_slot(cur, t);
      for (int k = 0; k < 2; ++k) {
        if (grp == k && has) {
          if (cur.s >= 0) add_slot(cur, t);
          // slots beyond the first kThreads of a large colour
          for (int s = co[c] + tg + kThreads; s < co[c + 1]; s += kThreads) {
            PrismIn in;
            load_slot(in, s);
            do_slot(in);
          }
        }
        __syncthreads();
      }
    }
  } else {
    for (int c = 0; c < kMaxColors; ++c) {
      const int s0 = co[c];
      if (s0 >= send) break;
      const int s1 = co[c + 1];
      PrismIn cur;
      load_slot(cur, first_slot(c));
      if (cur.s >= 0) do_slot(cur);
      // slots beyond the first kThreads of a large colour
      for (int s = s0 + tg + kThreads; s < s1; s += kThreads) {
        PrismIn in;
        load_slot(in, s);
        do_slot(in);
      }
      __syncthreads();
    }
  }
  for (int l = threadIdx.x; l < nown; l += blockDim.x) {
    if constexpr (ACC > 0)
      da[n0 + l] += acc[l];
    else
      da[n0 + l] = acc[l];
  }
}

}  // namespace bhip

extern "C" {

}  // extern "C"

namespace bhip {
// out[i] = in[perm[i] + shift] (0 past the end): a node field of the
// caller shifted by `shift` nodes, into the context's order
__global__ void k_gather_shift(double* __restrict__ out, const double* __restrict__ in,
                               const int32_t* __restrict__ perm, int shift, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int j = perm[i] + shift;
    out[i] = j < n ? in[j] : 0.0;
  }
}

// The operand-preparation entry points run with host arrays (synchronous:
// copied in, waited for) or, in their _dev forms, with device arrays in the
// caller's order (queued on the context stream only). Marks the call.
struct InputMode {
  bhip_ctx* c;
  InputMode(bhip_ctx* cc, bool dev) : c(cc) { c->in_dev = dev; }
  ~InputMode() { c->in_dev = false; }
};

// node field src (caller order, host or device per the call) into node
// scratch slot k in the context's order, shifted by `shift` caller nodes;
// returns the slot's address (nullptr on error)
static double* node_in(bhip_ctx* c, int k, const double* src, int shift = 0) {
  const int n = c->L.npoin;
  if (c->nsc.alloc((size_t)7 * n) != hipSuccess) return nullptr;
  double* dst = c->nsc.p + (size_t)k * n;
  const double* from = src;
  if (!c->in_dev) {
    double* stage = c->nsc.p + (size_t)6 * n;
    if (to_device(c, stage, src, sizeof(double) * n) !=
        hipSuccess)
      return nullptr;
    from = stage;
  }
  hipLaunchKernelGGL(k_gather_shift, dim3(c->grid_nodes), dim3(kThreads), 0, c->stream, dst, from,
                     c->perm.p, shift, n);
  if (hipGetLastError() != hipSuccess) return nullptr;
  // the staging area is reused by the next host field
  if (!c->in_dev && hipStreamSynchronize(c->stream) != hipSuccess) return nullptr;
  return dst;
}

// element field src (caller element order, read per slot through
// slot_elem): a device array is used as it is, a host one copied to element
// scratch slot k
static const double* elem_in(bhip_ctx* c, int k, const double* src) {
  if (c->in_dev) return src;
  const int ne = c->L.nelem;
  if (c->esc.alloc((size_t)4 * ne) != hipSuccess) return nullptr;
  double* dst = c->esc.p + (size_t)k * ne;
  if (to_device(c, dst, src, sizeof(double) * ne) !=
      hipSuccess)
    return nullptr;
  return dst;
}

// MT02PP's node record in the context's order: (Z, F, G, H + H_up), H_up
// being H at the caller's node NPOU0 above (mt02pp.f:233-235: II4 = I1 +
// NPOU0; zero past the last node, which is never a prism's lower node)
__global__ void k_gather_prism4(double* __restrict__ out, const double* __restrict__ z,
                                const double* __restrict__ f, const double* __restrict__ g,
                                const double* __restrict__ h, const int32_t* __restrict__ perm,
                                int npou0, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int j = perm[i], ju = j + npou0;
    const double hu = ju < n ? h[ju] : 0.0;
    double2* o = reinterpret_cast<double2*>(out + 4 * (size_t)i);
    o[0] = make_double2(z[j], f[j]);
    o[1] = make_double2(g[j], h[j] + hu);
  }
}

// the four MT02PP node fields (caller order, host or device per the call)
// into the node record array; returns its address (nullptr on error)
static double* node_in_prism4(bhip_ctx* c, const double* z, const double* f, const double* g,
                              const double* h, int npou0) {
  const int n = c->L.npoin;
  if (c->nsc.alloc((size_t)8 * n) != hipSuccess) return nullptr;
  const double* src[4] = {z, f, g, h};
  if (!c->in_dev) {
    for (int k = 0; k < 4; ++k) {
      double* stage = c->nsc.p + (size_t)(4 + k) * n;
      if (to_device(c, stage, src[k], sizeof(double) * n) !=
          hipSuccess)
        return nullptr;
      src[k] = stage;
    }
  }
  hipLaunchKernelGGL(k_gather_prism4, dim3(c->grid_nodes), dim3(kThreads), 0, c->stream,
                     c->nsc.p, src[0], src[1], src[2], src[3], c->perm.p, npou0, n);
  if (hipGetLastError() != hipSuccess) return nullptr;
  if (!c->in_dev && hipStreamSynchronize(c->stream) != hipSuccess) return nullptr;
  return c->nsc.p;
}

// the end of an operand-preparation call: host calls wait for their work
static int input_done(bhip_ctx* c) {
  if (!c->in_dev) HIP_OK(hipStreamSynchronize(c->stream));
  return BHIP_OK;
}

// DIRICH (dirich.f:83 -> diri01.f:94-139 -> diraux.f) on f, sm, fbor, limdir in
// the caller's order, host or device arrays per the call (InputMode)
static int dirich_impl(bhip_ctx* c, double* f, double* sm, const double* fbor,
                       const int32_t* limdir, int kdir) {
  const Layout& L = c->L;
  const int n = L.npoin;
  const int g = c->grid_nodes;
  const double* nb = node_in(c, 0, fbor);
  const double* ns = nb ? node_in(c, 1, sm) : nullptr;
  const double* nf = ns ? node_in(c, 2, f) : nullptr;
  if (!nf) return BHIP_EDEVICE;
  HIP_OK(c->isc.alloc((size_t)2 * n));
  int32_t* il = c->isc.p;
  const int32_t* lsrc = limdir;
  if (!c->in_dev) {
    HIP_OK(to_device(c, c->isc.p + n, limdir, sizeof(int32_t) * n));
    lsrc = c->isc.p + n;
  }
  // internal order: vio0 = FBOR, vd0 = SM, vd1 = F, il = LIMDIR
  HIP_OK(hipMemcpyAsync(c->vio0.p, nb, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
  HIP_OK(hipMemcpyAsync(c->vd0.p, ns, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
  HIP_OK(hipMemcpyAsync(c->vd1.p, nf, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
  hipLaunchKernelGGL(k_gather_any<int32_t>, dim3(g), dim3(kThreads), 0, c->stream, il, lsrc,
                     c->perm.p, n);
  HIP_OK(hipGetLastError());
  // WORK2 = A FBOR (diri01.f:103-110), SM -= WORK2, DIRAUX
  int rc = matvec_internal(c, BHIP_OP_AY, 0.0, c->vio0.p, c->vio0.p, c->vr.p);
  if (rc) return rc;
  hipLaunchKernelGGL(k_diraux, dim3(g), dim3(kThreads), 0, c->stream, c->vd0.p, c->vr.p,
                     c->da.p, c->vio0.p, il, kdir, c->vdinv.p, c->vd1.p, n);
  HIP_OK(hipGetLastError());
  // OM 'M=DMD' with W and the diagonal declared '0' (diri01.f:131-136)
  if (c->ncoef > 0) {
    HIP_OK(launch_dmd(c, (float*)nullptr, c->vdinv.p));
  }
  if (c->in_dev) {
    hipLaunchKernelGGL(k_scatter<double>, dim3(g), dim3(kThreads), 0, c->stream, sm, c->vd0.p,
                       c->perm.p, n);
    hipLaunchKernelGGL(k_scatter<double>, dim3(g), dim3(kThreads), 0, c->stream, f, c->vd1.p,
                       c->perm.p, n);
    HIP_OK(hipGetLastError());
    return BHIP_OK;
  }
  hipLaunchKernelGGL(k_scatter<double>, dim3(g), dim3(kThreads), 0, c->stream, c->vio1.p,
                     c->vd0.p, c->perm.p, n);
  HIP_OK(hipGetLastError());
  HIP_OK(to_host(c, sm, c->vio1.p, sizeof(double) * n));
  HIP_OK(hipStreamSynchronize(c->stream));
  hipLaunchKernelGGL(k_scatter<double>, dim3(g), dim3(kThreads), 0, c->stream, c->vio1.p,
                     c->vd1.p, c->perm.p, n);
  HIP_OK(hipGetLastError());
  HIP_OK(to_host(c, f, c->vio1.p, sizeof(double) * n));
  HIP_OK(hipStreamSynchronize(c->stream));
  return BHIP_OK;
}

// LUMP (lump.f): DIAG = A * XMUL, caller order
static int lump_impl(bhip_ctx* c, double xmul, double* diag) {
  const int n = c->L.npoin;
  const int g = c->grid_nodes;
  hipLaunchKernelGGL(k_fill<double>, dim3(g), dim3(kThreads), 0, c->stream, c->vio0.p, xmul,
                     (long)n);
  HIP_OK(hipGetLastError());
  int rc = matvec_internal(c, BHIP_OP_AY, 0.0, c->vio0.p, c->vio0.p, c->vr.p);
  if (rc) return rc;
  double* dst = c->in_dev ? diag : c->vio1.p;
  hipLaunchKernelGGL(k_scatter<double>, dim3(g), dim3(kThreads), 0, c->stream, dst, c->vr.p,
                     c->perm.p, n);
  HIP_OK(hipGetLastError());
  if (!c->in_dev) {
    HIP_OK(to_host(c, diag, c->vio1.p, sizeof(double) * n));
    HIP_OK(hipStreamSynchronize(c->stream));
  }
  return BHIP_OK;
}
}  // namespace bhip

extern "C" {

int bhip_dirich(bhip_ctx* c, double* f, double* sm, const double* fbor, const int32_t* limdir,
                int kdir) {
  if (!c || !c->have_matrix || !f || !sm || !fbor || !limdir) return BHIP_EBADARG;
  InputMode mode(c, false);
  return dirich_impl(c, f, sm, fbor, limdir, kdir);
}

int bhip_dirich_dev(bhip_ctx* c, double* f, double* sm, const double* fbor,
                    const int32_t* limdir, int kdir) {
  if (!c || !c->have_matrix || !f || !sm || !fbor || !limdir) return BHIP_EBADARG;
  InputMode mode(c, true);
  return dirich_impl(c, f, sm, fbor, limdir, kdir);
}

int bhip_lump(bhip_ctx* c, double xmul, double* diag) {
  if (!c || !c->have_matrix || !diag) return BHIP_EBADARG;
  InputMode mode(c, false);
  return lump_impl(c, xmul, diag);
}

int bhip_lump_dev(bhip_ctx* c, double xmul, double* diag) {
  if (!c || !c->have_matrix || !diag) return BHIP_EBADARG;
  InputMode mode(c, true);
  return lump_impl(c, xmul, diag);
}

}  // extern "C"

namespace bhip {
static int elem_common(bhip_ctx* c, int kind, double xmul, const double* xel, const double* yel,
                       const double* surfac, const double* u, const double* v,
                       int accumulate) {
  if (!c || c->L.ndp != 3) return BHIP_EBADARG;
  // XEL, YEL, SURFAC all NULL: the context's geometry (bhip_ctx_set_geometry)
  const bool use_geo = !surfac && !xel && !yel;
  if (use_geo ? !c->have_geo : (!surfac || (kind != 100 && (!xel || !yel))))
    return BHIP_EBADARG;
  if (kind != 100 && !u) return BHIP_EBADARG;
  if (kind == 2 && !v) return BHIP_EBADARG;
  const Layout& L = c->L;
  const int n = L.npoin, ne = L.nelem;
  if (accumulate && (!c->have_matrix || c->typdia != 'Q' || c->typext == '0'))
    return BHIP_EBADARG;
  const int target_q = accumulate && c->typext == 'Q';
  if (!c->have_matrix) {
    HIP_OK(c->da.alloc(n));
    HIP_OK(c->xa.alloc((size_t)6 * L.nslot_pad));
  }
  // the caller's geometry arrays (host calls only): temporaries
  DevBuf<double> dS, dX, dY;
  if (!use_geo) {
    HIP_OK(dS.alloc(ne));
    HIP_OK(to_device(c, dS.p, surfac, sizeof(double) * ne));
    if (kind != 100) {
      HIP_OK(dX.alloc((size_t)3 * ne));
      HIP_OK(dY.alloc((size_t)3 * ne));
      HIP_OK(to_device(c, dX.p, xel, sizeof(double) * 3 * ne));
      HIP_OK(to_device(c, dY.p, yel, sizeof(double) * 3 * ne));
    }
  }
  const double* dU = nullptr;
  const double* dV = nullptr;
  if (kind != 100) {
    if (!(dU = node_in(c, 0, u))) return BHIP_EDEVICE;
    if (v && !(dV = node_in(c, 1, v))) return BHIP_EDEVICE;
  }
  const size_t lds = (size_t)L.max_owned * sizeof(double);
  hipLaunchKernelGGL(k_elem_tri, dim3(L.nblocks), dim3(kThreads), lds, c->stream, kind, xmul,
                     dS.p, dX.p, dY.p, ne, dU, v ? dV : dU, c->slot_elem.p, c->lid.p,
                     L.nslot_pad, c->node_start.p, c->halo_start.p, c->halo.p, c->color_off.p,
                     L.nblocks, c->xa.p, c->da.p, accumulate, target_q,
                     use_geo ? c->geo.p : nullptr);
  HIP_OK(hipGetLastError());
  // host calls, and calls with temporaries to free, wait for the kernel
  if (!use_geo) HIP_OK(hipStreamSynchronize(c->stream));
  if (input_done(c)) return BHIP_EDEVICE;
  // MATRIY sets TYPDIA 'Q', TYPEXT 'S' (matriy.f:375-376); 'M=M+N' of two
  // symmetric matrices stays symmetric, into a 'Q' one stays 'Q'
  c->typdia = 'Q';
  c->typext = target_q ? 'Q' : 'S';
  c->ncoef = target_q ? 6 : 3;
  c->have_matrix = true;
  return BHIP_OK;
}

// prism element matrices (host inputs): kind 200 MT01PP, 201 MT02PP
static int elem_prism(bhip_ctx* c, int kind, double xmul, const double* xel, const double* yel,
                      const double* surfac, const double* z, const double* f, const double* g,
                      const double* h, int h_zero, int npou0, int inchyd, int mon,
                      int accumulate) {
  if (!c || !z || c->L.ndp != 6) return BHIP_EBADARG;
  // XEL, YEL, SURFAC all NULL: the context's geometry (bhip_ctx_set_geometry)
  const bool use_geo = !surfac && !xel && !yel;
  if (use_geo ? !c->have_geo : (!surfac || (kind == 201 && (!xel || !yel))))
    return BHIP_EBADARG;
  if (kind == 201 && (!f || !g || !h)) return BHIP_EBADARG;
  const Layout& L = c->L;
  const int n = L.npoin, ne = L.nelem;
  if (npou0 < 0 || npou0 >= n) return BHIP_EBADARG;
  if (accumulate && (!c->have_matrix || c->typdia != 'Q' || c->typext == '0'))
    return BHIP_EBADARG;
  const int target_q = accumulate && c->typext == 'Q';
  if (!c->have_matrix) {
    HIP_OK(c->da.alloc(n));
    HIP_OK(c->xa.alloc((size_t)30 * L.nslot_pad));
  }
  // the caller's geometry arrays (host calls only): temporaries
  DevBuf<double> dS, dX, dY;
  if (!use_geo) {
    HIP_OK(dS.alloc(ne));
    HIP_OK(to_device(c, dS.p, surfac, sizeof(double) * ne));
    if (kind == 201) {
      HIP_OK(dX.alloc((size_t)3 * ne));
      HIP_OK(dY.alloc((size_t)3 * ne));
      HIP_OK(to_device(c, dX.p, xel, sizeof(double) * 3 * ne));
      HIP_OK(to_device(c, dY.p, yel, sizeof(double) * 3 * ne));
    }
  }
  // node fields in the context's order: Z alone (MT01PP) or MT02PP's record
  const double* nd = kind == 200 ? node_in(c, 0, z) : node_in_prism4(c, z, f, g, h, npou0);
  if (!nd) return BHIP_EDEVICE;
  // LDS: the diagonal accumulator, then (when the block's records fit the
  // device's per-workgroup LDS) the records of its local nodes
  const int rec_off = (L.max_owned + 1) & ~1;
  const size_t lds_acc = (size_t)L.max_owned * sizeof(double);
  const size_t lds_rec =
      ((size_t)rec_off + (size_t)(kind == 200 ? 1 : 4) * L.max_local) * sizeof(double);
  int lds_max = 0;
  if (hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device) !=
      hipSuccess)
    return BHIP_EDEVICE;
  const bool staged = lds_rec <= (size_t)lds_max;
  auto launch = [&](auto kern, size_t lds, int threads) -> int {
    if (lds > 65536)
      HIP_OK(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(kern, dim3(L.nblocks), dim3(threads), lds, c->stream, xmul, dS.p, dX.p,
                       dY.p, ne, nd, h_zero, inchyd, mon, c->slot_elem.p, c->lid.p, L.nslot_pad,
                       c->node_start.p, c->halo_start.p, c->halo.p, c->color_off.p, L.nblocks,
                       c->xa.p, c->da.p, use_geo ? c->geo.p : nullptr, rec_off, (int)L.col_ids);
    return BHIP_OK;
  };
  // staged: two groups of kThreads per workgroup
  auto pick = [&](auto k200, auto k201, size_t lds, int threads) {
    return kind == 200 ? launch(k200, lds, threads) : launch(k201, lds, threads);
  };
  const int acc_kind = !accumulate ? 0 : target_q ? 2 : 1;
#define BHIP_PRISM_PICK(ST, ACC)                                                             \
  (ST ? pick(k_elem_prism<200, 2, ACC>, k_elem_prism<201, 2, ACC>, lds_rec, 2 * kThreads) \
      : pick(k_elem_prism<200, 0, ACC>, k_elem_prism<201, 0, ACC>, lds_acc, kThreads))
  const int ls = staged ? (acc_kind == 0   ? BHIP_PRISM_PICK(1, 0)
                           : acc_kind == 1 ? BHIP_PRISM_PICK(1, 1)
                                           : BHIP_PRISM_PICK(1, 2))
                        : (acc_kind == 0   ? BHIP_PRISM_PICK(0, 0)
                           : acc_kind == 1 ? BHIP_PRISM_PICK(0, 1)
                                           : BHIP_PRISM_PICK(0, 2));
#undef BHIP_PRISM_PICK
  if (ls != BHIP_OK) return ls;
  HIP_OK(hipGetLastError());
  if (!use_geo) HIP_OK(hipStreamSynchronize(c->stream));
  if (input_done(c)) return BHIP_EDEVICE;
  c->typdia = 'Q';
  c->typext = target_q ? 'Q' : 'S';
  c->ncoef = target_q ? 30 : 15;
  c->have_matrix = true;
  return BHIP_OK;
}

}  // namespace bhip

extern "C" {

int bhip_ctx_set_geometry(bhip_ctx* c, const double* xel, const double* yel,
                          const double* surfac) {
  if (!c || !xel || !yel || !surfac || (c->L.ndp != 3 && c->L.ndp != 6)) return BHIP_EBADARG;
  const Layout& L = c->L;
  const int ne = L.nelem;
  DevBuf<double> dS, dX, dY;
  HIP_OK(dS.alloc(ne));
  HIP_OK(dX.alloc((size_t)3 * ne));
  HIP_OK(dY.alloc((size_t)3 * ne));
  HIP_OK(c->geo.alloc((size_t)5 * L.nslot_pad));
  HIP_OK(to_device(c, dS.p, surfac, sizeof(double) * ne));
  HIP_OK(to_device(c, dX.p, xel, sizeof(double) * 3 * ne));
  HIP_OK(to_device(c, dY.p, yel, sizeof(double) * 3 * ne));
  hipLaunchKernelGGL(k_geo_slots, dim3(grid_for(L.nslot)), dim3(kThreads), 0, c->stream,
                     c->geo.p, dX.p, dY.p, dS.p, c->slot_elem.p, L.nslot, L.nslot_pad, ne);
  HIP_OK(hipGetLastError());
  HIP_OK(hipStreamSynchronize(c->stream));
  c->have_geo = true;
  return BHIP_OK;
}

int bhip_matrix_mt01pp(bhip_ctx* c, double xmul, const double* surfac, const double* z,
                       int accumulate) {
  return elem_prism(c, 200, xmul, nullptr, nullptr, surfac, z, nullptr, nullptr, nullptr, 0, 0,
                    0, 0, accumulate);
}

int bhip_matrix_mt02pp(bhip_ctx* c, double xmul, const double* xel, const double* yel,
                       const double* surfac, const double* z, const double* f, const double* g,
                       const double* h, int h_zero, int npou0, int inchyd, int mon,
                       int accumulate) {
  return elem_prism(c, 201, xmul, xel, yel, surfac, z, f, g, h, h_zero, npou0, inchyd, mon,
                    accumulate);
}

int bhip_matrix_mt01aa(bhip_ctx* c, double xmul, const double* surfac, int accumulate) {
  return elem_common(c, 100, xmul, nullptr, nullptr, surfac, nullptr, nullptr, accumulate);
}

int bhip_matrix_mt02aa(bhip_ctx* c, int kind, double xmul, const double* xel, const double* yel,
                       const double* surfac, const double* u, const double* v, int accumulate) {
  if (kind != 1 && kind != 2) return BHIP_EBADARG;
  return elem_common(c, kind, xmul, xel, yel, surfac, u, v, accumulate);
}

// the same element matrices from device node fields (caller order) and the
// context's geometry, queued on the context stream
int bhip_matrix_mt01aa_dev(bhip_ctx* c, double xmul, int accumulate) {
  if (!c) return BHIP_EBADARG;
  InputMode mode(c, true);
  return elem_common(c, 100, xmul, nullptr, nullptr, nullptr, nullptr, nullptr, accumulate);
}

int bhip_matrix_mt02aa_dev(bhip_ctx* c, int kind, double xmul, const double* u_dev,
                           const double* v_dev, int accumulate) {
  if (!c || (kind != 1 && kind != 2)) return BHIP_EBADARG;
  InputMode mode(c, true);
  return elem_common(c, kind, xmul, nullptr, nullptr, nullptr, u_dev, v_dev, accumulate);
}

int bhip_matrix_mt01pp_dev(bhip_ctx* c, double xmul, const double* z_dev, int accumulate) {
  if (!c) return BHIP_EBADARG;
  InputMode mode(c, true);
  return elem_prism(c, 200, xmul, nullptr, nullptr, nullptr, z_dev, nullptr, nullptr, nullptr, 0,
                    0, 0, 0, accumulate);
}

int bhip_matrix_mt02pp_dev(bhip_ctx* c, double xmul, const double* z_dev, const double* f_dev,
                           const double* g_dev, const double* h_dev, int h_zero, int npou0,
                           int inchyd, int mon, int accumulate) {
  if (!c) return BHIP_EBADARG;
  InputMode mode(c, true);
  return elem_prism(c, 201, xmul, nullptr, nullptr, nullptr, z_dev, f_dev, g_dev, h_dev, h_zero,
                    npou0, inchyd, mon, accumulate);
}

}  // extern "C"

#include "wave.inl"
#include "supg.inl"
