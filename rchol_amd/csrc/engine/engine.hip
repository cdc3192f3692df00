// rchol_amd engine: device residency, schedule analysis, PCG driver and the
// C ABI declared in include/rchol_amd.h.
//
// Reference path replaced (c++/util/pcg.cpp):
//   pcg::pcg / create_sparse      -> rchol_amd_create / upload_A / upload_G
//   pcg::iteration                -> Engine::pcg (device-resident loop)
//   pcg::matrix_vector_product    -> k_spmv_p / k_spmv_res
//   pcg::precond_solve            -> two k_sptrsv sweeps
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include <sys/mman.h>

#include "../../../include/rchol_amd.h"
#include "kernels.hpp"
#include "kernels_mr.hpp"
#include "kernels_seg.hpp"
#include "kernels_chunk.hpp"

using namespace rchol_amd;

// Scratch of the allocation-free preconditioner entry point
// (rchol_amd_precond_device_noalloc): a control block followed by the sweep
// vectors (layout in scratch_layout()).  Each GPU holds one copy for the life
// of the process.  Host code never passes the array's name to a kernel or a
// copy: its address on a device is looked up once, at the first
// rchol_amd_create on that device, and that address is used from then on.
constexpr size_t kScratchBytes = 64ull << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

// The error flag of a sweep as a message: bit 0 a dependency wait hit its
// bound, bit 1 a schedule held a task kind the kernel does not run, bit 2
// the persistent grid's workgroups were not all resident at once.
int stall(int err) {
  return fail(RCHOL_AMD_ESTALL, "triangular sweep stalled (err=%d:%s%s%s)", err,
              (err & 1) ? " a dependency wait exceeded its bound" : "",
              (err & 2) ? " a task kind outside its schedule" : "",
              (err & 4) ? " not every workgroup of the persistent sweep was resident" : "");
}

#define HIP_TRY(expr)                                                                    \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      return fail(RCHOL_AMD_EHIP, "%s failed: %s", #expr, hipGetErrorString(e_));        \
  } while (0)

constexpr int kTaskRows = 64;     // slots per triangular-sweep task
// Grid caps, in blocks per CU, of the SpMV and of the streaming vector
// kernels (measured at 128^3 and 256^3): a smaller SpMV grid keeps the rows
// an XCD works on at once -- and the vector lines they gather -- within its
// L2 (256^3: 0.86 ms at 8 per CU, 0.72 ms at 2, 1.02 ms at 1); the
// streaming kernels reach the copy rate with 4.
#ifndef RCHOL_AMD_SPMV_PER_CU
#define RCHOL_AMD_SPMV_PER_CU 2
#endif
constexpr int kSpmvBlocksPerCu = RCHOL_AMD_SPMV_PER_CU;
// The K-column SpMV gathers K values per entry and needs more rows in flight
// (K = 32 at 256^3, ms per SpMV: 4 per CU 12.7, 5 11.5, 6 10.9, 7 11.1,
// 8 12.1).
#ifndef RCHOL_AMD_SPMV_MR_PER_CU
#define RCHOL_AMD_SPMV_MR_PER_CU 6
#endif
constexpr int kSpmvMrBlocksPerCu = RCHOL_AMD_SPMV_MR_PER_CU;
#ifndef RCHOL_AMD_VEC_PER_CU
#define RCHOL_AMD_VEC_PER_CU 4
#endif
constexpr int kVecBlocksPerCu = RCHOL_AMD_VEC_PER_CU;
constexpr int kBlock = 256;       // threads per block for every kernel
constexpr size_t kIdxMax = 0x7FFFFFFFu;
constexpr size_t kScratchCtl = 256;   // scratch: DevScalars of the standalone sweeps
constexpr int kMaxDevices = 64;
static_assert(sizeof(DevScalars) <= kScratchCtl, "control block outgrew its slot");

template <typename T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { reset(); }
  void reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  hipError_t alloc(size_t count) {
    reset();
    n = count;
    return hipMalloc((void **)&p, std::max<size_t>(count, 1) * sizeof(T));
  }
  hipError_t upload(const T *h, size_t count) {
    hipError_t e = alloc(count);
    if (e != hipSuccess) return e;
    return count ? hipMemcpy(p, h, count * sizeof(T), hipMemcpyHostToDevice) : hipSuccess;
  }
};

// One triangular sweep on the device: rows in level ("slot") order.
struct Sweep {
  int nslots = 0, ntasks = 0, nlevels = 0;
  size_t nnz_off = 0;
  DevBuf<int> task_ptr, slot_row, rp, ci;
  // split schedule: levels [0, head_end) and [tail_begin, nlevels) run in one
  // workgroup (k_sptrsv_serial), the middle in the persistent kernel
  std::vector<int> level_ptr_h, rp_h;
  std::vector<unsigned short> proxy_h;  // per slot: proxy distance (0: none)
  DevBuf<int> level_ptr, level_lg;
  DevBuf<int> wave_ptr, wt_s0, wt_info;  // middle tasks in per-wave lists
  int head_end = 0, tail_begin = 0, mid_ntasks = 0;
  DevBuf<double> cv64, diag64;
  DevBuf<float> cv32, diag32;
  // chunked form (k_sptrsv_chunk), built when the factor's chains are deep
  struct Chunks {
    int nchunks = 0, nparts = 0;
    long long rounds = 0;
    double est_us = 0.0;
    DevBuf<int> order, q0, m, r0, r1, split, q_row, x_l, x_ci, dep_ptr, dep;
    DevBuf<int> c_part0, p_chunk, p_l0, p_l1, p_lg, x_rp;  // external parts
    DevBuf<double> ext64;  // the parts' external sums (sweep precision)
    DevBuf<float> ext32;
    DevBuf<int> done;  // per chunk: published (reset by each launch's last workgroup)
    DevBuf<unsigned> r_w;
    DevBuf<double> q_diag64, x_cv64, r_v64;
    DevBuf<float> q_diag32, x_cv32, r_v32;
    void reset() {
      nchunks = 0;
      nparts = 0;
      rounds = 0;
      est_us = 0.0;
      for (DevBuf<int> *b : {&order, &q0, &m, &r0, &r1, &split, &q_row, &x_l, &x_ci, &dep_ptr, &dep,
                             &done, &c_part0, &p_chunk, &p_l0, &p_l1, &p_lg, &x_rp})
        b->reset();
      ext64.reset();
      ext32.reset();
      r_w.reset();
      q_diag64.reset();
      x_cv64.reset();
      r_v64.reset();
      q_diag32.reset();
      x_cv32.reset();
      r_v32.reset();
    }
  } ch;
};

#include "analysis.hpp"

#include "chunk_plan.hpp"

// Cut a sweep's levels into a narrow head, a middle and a narrow tail
// (levels of at most `width` rows run in one workgroup; width 0 = no split).
void compute_split(const std::vector<int> &lp, int L, int width, int &h, int &t) {
  h = 0;
  t = L;
  if (width > 0) {
    while (h < L && lp[h + 1] - lp[h] <= width) h++;
    while (t > h && lp[t] - lp[t - 1] <= width) t--;
  }
}

// Lane width (log2) for a row of len entries: the smallest width >= the
// minimum that covers the row in one round of kGather gathers per lane.
int lane_lg(int len, int lmin_lg) {
  int lg = lmin_lg;
  while (lg < 6 && (kGather << lg) < len) lg++;
  return lg;
}

// Tasks of the middle levels [h, t), in level order and, inside a level,
// in slot (nested-dissection) order.  A level is wide when one-pass tasks
// of 16 rows would outnumber the waves: its slots go to stream tasks of up
// to 64 rows and 64 * kStream entries (one round of gathers), or, when its
// rows read nothing (level 0 of the forward sweep: flat_ok), to flat tasks
// of 64 * kFlatPerLane.  A narrow
// level is cut into one-pass lane tasks whose width L covers the longest of
// its 64/L rows in one round of kGather gathers per lane, so it spreads
// over as many waves as it can use.
#ifndef RCHOL_AMD_WIDE_ROWS
#define RCHOL_AMD_WIDE_ROWS 16
#endif
constexpr long long kWideRowsPerWave = RCHOL_AMD_WIDE_ROWS;  // stream-task threshold

struct TaskCut {
  std::vector<int> s0, info, level;  // info = cnt | lg << kTaskInfoShift
};

void cut_tasks(const std::vector<int> &lp, const std::vector<int> &rp,
               const std::vector<unsigned short> &proxy, int h, int t, int lmin_lg,
               long long nwaves, bool flat_ok, TaskCut &tc) {
  tc.s0.clear();
  tc.info.clear();
  tc.level.clear();
  auto push = [&](int s, int cnt, int lg, int l) {
    // a lane task's proxy: that of its first row that has one in the first
    // round (one pass: row k is row group k)
    int px = 0;
    if (RCHOL_AMD_PROXY && lg < kTaskStream && (size_t)(s + cnt) <= proxy.size())
      for (int k = 0; k < cnt; k++) {
        const int pc = proxy[s + k];
        if (pc > 0 && pc < (kGather << lg)) {
          px = pc | (k << 9);
          break;
        }
      }
    tc.s0.push_back(s);
    tc.info.push_back(cnt | (lg << kTaskInfoShift) | (px << kTaskProxyShift));
    tc.level.push_back(l);
  };
  for (int l = h; l < t; l++) {
    const int s_end = lp[l + 1];
    const long long rows = s_end - lp[l];
    const bool wide = rows > kWideRowsPerWave * nwaves;
    if (flat_ok && wide && rp[s_end] == rp[lp[l]]) {  // a wide level that reads nothing
      for (int s = lp[l]; s < s_end; s += 64 * kFlatPerLane)
        push(s, std::min(s_end - s, 64 * kFlatPerLane), kTaskFlat, l);
      continue;
    }
    for (int s = lp[l]; s < s_end;) {
      if (wide) {
        int k = 1;  // at least one row, however long (the kernel loops rounds)
        while (s + k < s_end && k < 64 && rp[s + k + 1] - rp[s] <= 64 * kStream) k++;
        push(s, k, kTaskStream, l);
        s += k;
        continue;
      }
      int lg = lane_lg(rp[s + 1] - rp[s], lmin_lg);
      for (;;) {  // widen until the longest of the 64/L rows fits one round
        const int take = std::min(s_end - s, 64 >> lg);
        int need = lg;
        for (int k = 0; k < take; k++)
          need = std::max(need, lane_lg(rp[s + k + 1] - rp[s + k], lmin_lg));
        if (need == lg) break;
        lg = need;
      }
      const int take = std::min(s_end - s, 64 >> lg);
      push(s, take, lg, l);
      s += take;
    }
  }
}

// Tasks for the K-column sweeps: every level cut into runs of rows with an
// entry-lane width Le = 2^lg <= 64/K (a row takes K * Le lanes) that covers
// the longest of the run's rows in one round of kGather gathers, R = 64/(K Le)
// rows per pass; narrow levels get one-pass tasks, wide ones as many passes
// (at most 64 slots) as it takes to need about one task per wave.
void cut_tasks_mr(const std::vector<int> &lp, const std::vector<int> &rp, int nlevels, int K,
                  long long nwaves, TaskCut &tc) {
  tc.s0.clear();
  tc.info.clear();
  tc.level.clear();
  int maxlg = 0;
  while ((K << (maxlg + 1)) <= 64) maxlg++;
  for (int l = 0; l < nlevels; l++) {
    const int s_end = lp[l + 1];
    const long long rows = s_end - lp[l];
    for (int s = lp[l]; s < s_end;) {
      int lg = std::min(lane_lg(rp[s + 1] - rp[s], 0), maxlg), take = 0;
      for (;;) {
        const int R = 64 / (K << lg);
        const long long one_pass = (rows + R - 1) / R;
        const int passes =
            (int)std::min<long long>(64 / R, std::max<long long>(1, (one_pass + nwaves - 1) / nwaves));
        take = std::min(s_end - s, passes * R);
        int need = lg;
        for (int k = 0; k < take; k++)
          need = std::max(need, std::min(lane_lg(rp[s + k + 1] - rp[s + k], 0), maxlg));
        if (need == lg) break;
        lg = need;
      }
      tc.s0.push_back(s);
      tc.info.push_back(take | (lg << kTaskInfoShift));
      tc.level.push_back(l);
      s += take;
    }
  }
}

// Static task -> wave assignment.  Block b runs on XCD b % 8 (round-robin
// dispatch), so a level's tasks are split into 8 contiguous runs, run x
// going to the waves of XCD x: neighbouring rows, and the rows of the same
// piece of the domain in consecutive levels, share one L2.  Inside an XCD
// the waves take tasks round robin with a cursor that carries across
// levels.  Each wave's list is in level order, which is all the
// deadlock-freedom argument of k_sptrsv needs.
void assign_waves(const TaskCut &tc, int nblocks, std::vector<int> &wave_ptr,
                  std::vector<int> &ws0, std::vector<int> &winfo) {
  const int W = nblocks * 4;
  const int X = nblocks >= 8 ? 8 : 1;
  std::vector<std::vector<int>> xw(X);  // waves of each XCD
  for (int b = 0; b < nblocks; b++)
    for (int q = 0; q < 4; q++) xw[b % X].push_back(4 * b + q);
  std::vector<size_t> cursor(X, 0);
  std::vector<std::vector<int>> lists(W);
  const size_t T = tc.s0.size();
  for (size_t i = 0; i < T;) {
    size_t j = i;
    while (j < T && tc.level[j] == tc.level[i]) j++;
    const size_t tl = j - i;
    for (size_t k = 0; k < tl; k++) {
      const int x = (int)((k * X) / tl);
      const int w = xw[x][cursor[x]++ % xw[x].size()];
      lists[w].push_back((int)(i + k));
    }
    i = j;
  }
  wave_ptr.assign(W + 1, 0);
  ws0.clear();
  winfo.clear();
  for (int w = 0; w < W; w++) {
    for (int k : lists[w]) {
      ws0.push_back(tc.s0[k]);
      winfo.push_back(tc.info[k]);
    }
    wave_ptr[w + 1] = (int)ws0.size();
  }
  if (ws0.empty()) {
    ws0.push_back(0);
    winfo.push_back(0);
  }
}

// Upload a sweep's split schedule: head/tail bounds [0, h) / [t, nlevels),
// per-level lane widths (for the one-workgroup phases) and the middle's
// tasks listed per wave of a grid of nblocks persistent blocks.
hipError_t upload_schedule(Sweep &sw, int h, int t, int lmin_lg, const TaskCut &tc, int nblocks) {
  const std::vector<int> &lp = sw.level_ptr_h;
  std::vector<int> llg(std::max(sw.nlevels, 1), lmin_lg);
  for (int l = 0; l < sw.nlevels; l++) {
    int lg = lmin_lg;
    for (int k = lp[l]; k < lp[l + 1]; k++)
      lg = std::max(lg, lane_lg(sw.rp_h[k + 1] - sw.rp_h[k], lmin_lg));
    llg[l] = lg;
  }
  sw.head_end = h;
  sw.tail_begin = t;
  sw.mid_ntasks = (int)tc.s0.size();
  std::vector<int> wp, ws0, winfo;
  assign_waves(tc, nblocks, wp, ws0, winfo);
  hipError_t e = sw.level_lg.upload(llg.data(), llg.size());
  if (e == hipSuccess) e = sw.wave_ptr.upload(wp.data(), wp.size());
  if (e == hipSuccess) e = sw.wt_s0.upload(ws0.data(), ws0.size());
  if (e == hipSuccess) e = sw.wt_info.upload(winfo.data(), winfo.size());
  return e;
}

}  // namespace

struct rchol_amd_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int ncu = 0;
  size_t n = 0;
  // A in CSR, int32 indices
  bool hasA = false;
  size_t nnzA = 0;
  DevBuf<int> arp, aci;
  DevBuf<double> av;
  // G
  bool hasG = false;
  size_t nG = 0, nnzG = 0;
  int gprec = RCHOL_AMD_FP64;
  Sweep fwd, bwd;
  int trsv_grid = 0;
  // optional ordering: vectors handed to the solves are in the original
  // numbering, A and G in the permuted one (pcg runs on A(P,P))
  bool hasPerm = false;
  DevBuf<int> perm;
  DevBuf<double> pin, pout;  // gathered input / permuted output staging
  // work vectors
  size_t nwork = 0;
  DevBuf<double> b, x, r, q, p0, p1, partials, tmp64a;
  DevBuf<double> y64, z64;
  DevBuf<float> y32, z32;
  DevScalars *sc = nullptr;
  DevScalars *hsc = nullptr;  // pinned mirror
  // options
  int graph_iters = 16;
  int resid_period = 0;
  int spmv_lanes = 0;  // lanes per SpMV row; 0 = stream form (64 rows per wave step)
  int trsv_lanes = 1;  // minimum lanes per sweep row (tasks widen to fit)
  int profile = 0;
  int profile_no_graph = 0;
  int serial_width = 0;  // narrow-level threshold of the split sweeps (0 = off)
  int coherent_gather = 0;  // sweeps read x with agent-scope loads from the start
  int task_waves = 0;       // waves the task cut plans for (0 = resident capacity)
  int strict_sync = 0;      // grid reductions with release/acquire fences
  int chunk_mode = -1;      // RCHOL_AMD_OPT_CHUNKED
  bool chunk_auto = false;  // the upload's own choice (chunk_mode -1)
  bool use_chunks = false;  // single-RHS sweeps run k_sptrsv_chunk
  int lds_max = 0;          // LDS bytes a workgroup may allocate
  int chunk_grid = 0;       // workgroups of the chunked sweeps
  size_t chunk_lds = 0;     // their dynamic LDS bytes
  long long g_gen = 0;      // bumped by every upload_G (keys the batched schedule)
  std::vector<hipEvent_t> pev;  // 5 per captured iteration when profiling
  // captured iteration graph
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  rchol_amd_stats stats{};
  // multi-RHS work space (K interleaved columns)
  struct Batch {
    int K = 0;
    size_t n = 0;
    DevBuf<double> b, x, r, q, p0, p1, res, partials;
    DevBuf<double> y64, z64;
    DevBuf<float> y32, z32;
    BatchScalars *bs = nullptr, *hbs = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int grid = 0;
    // sweep schedule cut for K columns: per-wave task lists, [0] fwd, [1] bwd
    DevBuf<int> wave_ptr[2], wt_s0[2], wt_info[2];
    int sched_K = 0, ntasks[2] = {0, 0};
    long long sched_gen = -1;
    void drop_graph() {
      if (gexec) (void)hipGraphExecDestroy(gexec);
      if (graph) (void)hipGraphDestroy(graph);
      gexec = nullptr;
      graph = nullptr;
    }
  } batch;
  // uneven batch of stacked systems
  struct Segments {
    int S = 0, nblocks = 0;
    std::vector<size_t> off;
    DevBuf<int> seg, r0, r1, sb;
    SegScalars *ss = nullptr, *hss = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    void drop_graph() {
      if (gexec) (void)hipGraphExecDestroy(gexec);
      if (graph) (void)hipGraphDestroy(graph);
      gexec = nullptr;
      graph = nullptr;
    }
  } segs;

  ~rchol_amd_ctx() {
    segs.drop_graph();
    if (segs.ss) (void)hipFree(segs.ss);
    if (segs.hss) (void)hipHostFree(segs.hss);
    batch.drop_graph();
    if (batch.bs) (void)hipFree(batch.bs);
    if (batch.hbs) (void)hipHostFree(batch.hbs);
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    if (sc) (void)hipFree(sc);
    if (hsc) (void)hipHostFree(hsc);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    for (hipEvent_t ev : pev) (void)hipEventDestroy(ev);
    if (stream) (void)hipStreamDestroy(stream);
  }

  void drop_graph() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    gexec = nullptr;
    graph = nullptr;
    batch.drop_graph();
    segs.drop_graph();
  }

  int vec_grid() const {
    size_t g = (n + kBlock - 1) / kBlock;
    return (int)std::max<size_t>(1, std::min<size_t>(g, (size_t)ncu * kVecBlocksPerCu));
  }
  int spmv_grid() const {
    size_t rows_per_block = (kBlock / 64) * (spmv_lanes ? 64 / spmv_lanes : 64);
    size_t g = (n + rows_per_block - 1) / rows_per_block;
    g = std::max<size_t>(1, std::min<size_t>(g, (size_t)ncu * kSpmvBlocksPerCu));
    if (g > 8) g = (g + 7) / 8 * 8;  // whole XCD rounds for the row walk of the stream form
    return (int)g;
  }
  int ensure_work();
  int size_trsv_grid();
  int launch_iteration(int parity, int slot);
  int pcg(const double *d_b, double tol, int maxit, double *d_x_out, int keep_x, double *relres,
          int *itr);
  int precond(const double *d_r, double *d_z);
  // multi-RHS
  int ensure_batch(int K);
  int ensure_batch_schedule(int K, int per_cu);
  template <int K>
  int pcg_batch(const double *d_b, int ncols, double tol, int maxit, double *d_x, int keep_x,
                double *relres, int *itr);
  template <int K>
  int precond_batch(const double *d_r, double *d_z);
  template <int K>
  void launch_iteration_mr(int parity, int slot);
  template <int K>
  void launch_sweeps_mr(const double *rhs, bool with_dot, const int *stop, hipEvent_t *ev);
  // uneven batch
  int pcg_segments(const double *d_b, double tol, int maxit, double *d_x, int keep_x,
                   double *relres, int *itr);
  void launch_iteration_seg(int parity, int slot);
};

namespace {

// Launch the forward + backward sweeps.  pcg_mode: fused sentinel/dot
// epilogues for the PCG loop; otherwise a standalone apply z = M^-1 r.
// One sweep in up to three launches: serial head, persistent middle, serial
// tail (the split is empty unless RCHOL_AMD_OPT_SERIAL_WIDTH is set).
// FUSE 1: sentinel into aux; 2: r.z -> beta; 0: plain.
template <typename VT, typename XT, typename BT, int FUSE>
void launch_one_sweep(rchol_amd_ctx *c, const Sweep &sw, const VT *cv, const VT *dg,
                      const BT *rhs, XT *x, XT *aux, const double *dotv, const int *stop,
                      DevScalars *sc) {
  if (c->use_chunks && sw.ch.nchunks > 0) {
    const Sweep::Chunks &h = sw.ch;
    const VT *qd, *xcv, *rv;
    XT *ext;
    if constexpr (sizeof(VT) == 8) {
      qd = h.q_diag64.p; xcv = h.x_cv64.p; rv = h.r_v64.p;
    } else {
      qd = h.q_diag32.p; xcv = h.x_cv32.p; rv = h.r_v32.p;
    }
    if constexpr (sizeof(XT) == 8) ext = (XT *)h.ext64.p; else ext = (XT *)h.ext32.p;
    hipLaunchKernelGGL((k_sptrsv_chunk<VT, XT, BT, FUSE>), dim3(c->chunk_grid), dim3(kChunkThreads),
                       c->chunk_lds, c->stream, h.nchunks, h.nchunks + h.nparts, h.order.p, h.q0.p,
                       h.m.p, h.r0.p, h.r1.p, h.split.p, h.c_part0.p, h.p_chunk.p, h.p_l0.p,
                       h.p_l1.p, h.p_lg.p, h.x_rp.p, h.dep_ptr.p, h.dep.p, h.done.p, ext, h.q_row.p, qd,
                       h.x_l.p, h.x_ci.p, xcv, h.r_w.p, rv, rhs, x, aux, dotv, c->partials.p, sc,
                       stop);
    return;
  }
  const bool head = sw.head_end > 0, tail = sw.tail_begin < sw.nlevels;
  if (head)
    hipLaunchKernelGGL((k_sptrsv_serial<VT, XT, BT, FUSE>), dim3(1), dim3(1024), 0, c->stream, 0,
                       sw.head_end, sw.level_ptr.p, sw.level_lg.p, sw.slot_row.p, sw.rp.p,
                       sw.ci.p, cv, dg, rhs, x, aux, dotv, sc, stop, 0);
  if (sw.mid_ntasks > 0 || (FUSE == 2 && !tail))
    hipLaunchKernelGGL((k_sptrsv<VT, XT, BT, FUSE>), dim3(c->trsv_grid), dim3(kBlock), 0,
                       c->stream, sw.wave_ptr.p, sw.wt_s0.p, sw.wt_info.p,
                       sw.slot_row.p, sw.rp.p, sw.ci.p, cv, dg, rhs, x, aux, dotv, c->partials.p,
                       sc, stop, tail ? 1 : 0, c->coherent_gather);
  if (tail)
    hipLaunchKernelGGL((k_sptrsv_serial<VT, XT, BT, FUSE>), dim3(1), dim3(1024), 0, c->stream,
                       sw.tail_begin, sw.nlevels, sw.level_ptr.p, sw.level_lg.p, sw.slot_row.p,
                       sw.rp.p, sw.ci.p, cv, dg, rhs, x, aux, dotv, sc, stop, 1);
}

template <typename VT, typename XT>
int dispatch_sweeps(rchol_amd_ctx *c, const double *rhs, XT *y, XT *z, const int *stop,
                    bool with_dot, const double *dotv, hipEvent_t *ev = nullptr,
                    DevScalars *sc = nullptr) {
  if (!sc) sc = c->sc;
  const VT *fcv, *fdg, *bcv, *bdg;
  if constexpr (sizeof(VT) == 8) {
    fcv = c->fwd.cv64.p; fdg = c->fwd.diag64.p; bcv = c->bwd.cv64.p; bdg = c->bwd.diag64.p;
  } else {
    fcv = c->fwd.cv32.p; fdg = c->fwd.diag32.p; bcv = c->bwd.cv32.p; bdg = c->bwd.diag32.p;
  }
  launch_one_sweep<VT, XT, double, 1>(c, c->fwd, fcv, fdg, rhs, y, z, nullptr, stop, sc);
  if (ev) (void)hipEventRecord(ev[1], c->stream);
  if (with_dot)
    launch_one_sweep<VT, XT, XT, 2>(c, c->bwd, bcv, bdg, (const XT *)y, z, (XT *)nullptr, dotv,
                                    stop, sc);
  else
    launch_one_sweep<VT, XT, XT, 0>(c, c->bwd, bcv, bdg, (const XT *)y, z, (XT *)nullptr, nullptr,
                                    stop, sc);
  if (ev) (void)hipEventRecord(ev[2], c->stream);
  return RCHOL_AMD_OK;
}

template <typename ZT, int L>
void launch_spmv_p_t(rchol_amd_ctx *c, const ZT *z, const double *p_old, double *p_new) {
  hipLaunchKernelGGL((k_spmv_p<ZT, L>), dim3(c->spmv_grid()), dim3(kBlock), 0, c->stream,
                     (int)c->n, c->arp.p, c->aci.p, c->av.p, z, p_old, p_new, c->q.p, c->r.p,
                     c->partials.p, c->sc);
}

template <typename ZT>
void launch_spmv_p(rchol_amd_ctx *c, const ZT *z, const double *p_old, double *p_new) {
  switch (c->spmv_lanes) {
    case 0:
      hipLaunchKernelGGL((k_p_update<ZT>), dim3(c->vec_grid()), dim3(kBlock), 0, c->stream,
                         (int)c->n, z, p_old, p_new, c->sc);
      hipLaunchKernelGGL(k_spmv_p_s, dim3(c->spmv_grid()), dim3(kBlock), 0, c->stream, (int)c->n,
                         c->arp.p, c->aci.p, c->av.p, (const double *)p_new, c->q.p, c->r.p,
                         c->partials.p, c->sc);
      break;
    case 4: launch_spmv_p_t<ZT, 4>(c, z, p_old, p_new); break;
    case 16: launch_spmv_p_t<ZT, 16>(c, z, p_old, p_new); break;
    case 32: launch_spmv_p_t<ZT, 32>(c, z, p_old, p_new); break;
    default: launch_spmv_p_t<ZT, 8>(c, z, p_old, p_new); break;
  }
}

template <int L>
void launch_spmv_res_t(rchol_amd_ctx *c, const double *x, const double *b, double *y) {
  hipLaunchKernelGGL((k_spmv_res<L>), dim3(c->spmv_grid()), dim3(kBlock), 0, c->stream,
                     (int)c->n, c->arp.p, c->aci.p, c->av.p, x, b, y, c->partials.p, c->sc);
}

template <int L>
void launch_resid_replace_t(rchol_amd_ctx *c) {
  hipLaunchKernelGGL((k_resid_replace<L>), dim3(c->spmv_grid()), dim3(kBlock), 0, c->stream,
                     (int)c->n, c->arp.p, c->aci.p, c->av.p, (const double *)c->x.p,
                     (const double *)c->b.p, c->r.p, c->partials.p, c->sc);
}

void launch_resid_replace(rchol_amd_ctx *c) {
  switch (c->spmv_lanes) {
    case 0:
      hipLaunchKernelGGL(k_resid_replace_s, dim3(c->spmv_grid()), dim3(kBlock), 0, c->stream,
                         (int)c->n, c->arp.p, c->aci.p, c->av.p, (const double *)c->x.p,
                         (const double *)c->b.p, c->r.p, c->partials.p, c->sc);
      break;
    case 4: launch_resid_replace_t<4>(c); break;
    case 16: launch_resid_replace_t<16>(c); break;
    case 32: launch_resid_replace_t<32>(c); break;
    default: launch_resid_replace_t<8>(c); break;
  }
}

void launch_spmv_res(rchol_amd_ctx *c, const double *x, const double *b, double *y) {
  switch (c->spmv_lanes) {
    case 0:
      hipLaunchKernelGGL(k_spmv_res_s, dim3(c->spmv_grid()), dim3(kBlock), 0, c->stream, (int)c->n,
                         c->arp.p, c->aci.p, c->av.p, x, b, y, c->partials.p, c->sc);
      break;
    case 4: launch_spmv_res_t<4>(c, x, b, y); break;
    case 16: launch_spmv_res_t<16>(c, x, b, y); break;
    case 32: launch_spmv_res_t<32>(c, x, b, y); break;
    default: launch_spmv_res_t<8>(c, x, b, y); break;
  }
}

}  // namespace

// Workgroups per CU a persistent grid may count on being co-resident: the
// occupancy query less one.  The query packs register files exactly, while
// the allocation granules of the hardware (SGPR blocks in particular) can
// leave one workgroup per CU short of it; a persistent sweep whose grid is
// not fully resident waits on tasks that never start.
static int resident_blocks(int query) { return std::max(1, std::min(query, 8) - 1); }

// Persistent sweep grid: every block must be resident at once (the
// deadlock-freedom argument of k_sptrsv).  The occupancy is the minimum
// over the three epilogue variants.
template <typename VT, typename XT>
static int trsv_occupancy() {
  int a = 0, b = 0, c = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, (const void *)k_sptrsv<VT, XT, double, 1>,
                                                   kBlock, 0) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, (const void *)k_sptrsv<VT, XT, XT, 2>,
                                                   kBlock, 0) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&c, (const void *)k_sptrsv<VT, XT, XT, 0>,
                                                   kBlock, 0) != hipSuccess)
    return 0;
  return std::min(a, std::min(b, c));
}

// Grid of the persistent sweeps and the task cut that goes with it: the
// resident capacity fixes how finely narrow levels are split.
int rchol_amd_ctx::size_trsv_grid() {
  int per_cu = gprec == RCHOL_AMD_FP32 ? trsv_occupancy<float, float>()
                                       : trsv_occupancy<double, double>();
  if (per_cu <= 0) return fail(RCHOL_AMD_EHIP, "occupancy query for the sweep kernels failed");
  per_cu = resident_blocks(per_cu);
  const int cap_blocks = per_cu * ncu;
  const long long plan_waves = task_waves > 0 ? task_waves : (long long)cap_blocks * (kBlock / 64);
  int lmin_lg = 0;
  while ((1 << lmin_lg) < trsv_lanes) lmin_lg++;
  int h[2], t[2];
  TaskCut tc[2];
  Sweep *sw[2] = {&fwd, &bwd};
  size_t max_tasks = 0;
  for (int k = 0; k < 2; k++) {
    compute_split(sw[k]->level_ptr_h, sw[k]->nlevels, serial_width, h[k], t[k]);
    cut_tasks(sw[k]->level_ptr_h, sw[k]->rp_h, sw[k]->proxy_h, h[k], t[k], lmin_lg, plan_waves,
              k == 0, tc[k]);
    max_tasks = std::max(max_tasks, tc[k].s0.size());
  }
  trsv_grid = (int)std::max<size_t>(1, std::min<size_t>(cap_blocks, (max_tasks + 3) / 4));
  // the residency check counts arrivals in whole grids: restart the count
  HIP_TRY(hipMemset(&sc->arrive, 0, sizeof(sc->arrive)));
  for (int k = 0; k < 2; k++) HIP_TRY(upload_schedule(*sw[k], h[k], t[k], lmin_lg, tc[k], trsv_grid));
  stats.trsv_waves = trsv_grid * (kBlock / 64);
  stats.fwd_tasks = fwd.mid_ntasks;
  stats.bwd_tasks = bwd.mid_ntasks;
  return RCHOL_AMD_OK;
}

int rchol_amd_ctx::ensure_work() {
  if (nwork == n && b.p) return RCHOL_AMD_OK;
  drop_graph();
  HIP_TRY(b.alloc(n));
  HIP_TRY(x.alloc(n));
  HIP_TRY(r.alloc(n));
  HIP_TRY(q.alloc(n));
  HIP_TRY(p0.alloc(n));
  HIP_TRY(p1.alloc(n));
  HIP_TRY(tmp64a.alloc(n));
  HIP_TRY(y64.alloc(n));
  HIP_TRY(z64.alloc(n));
  HIP_TRY(y32.alloc(n));
  HIP_TRY(z32.alloc(n));
  HIP_TRY(partials.alloc(std::max<size_t>((size_t)ncu * 8 * 4 + 64, (size_t)segs.nblocks * 2 + 64)));
  nwork = n;
  return RCHOL_AMD_OK;
}

// One PCG iteration (pcg.cpp:82-110 body): M^-1 r, p update + A p + alpha,
// x/r update + loop test.  parity selects the p ping-pong buffers.
int rchol_amd_ctx::launch_iteration(int parity, int slot) {
  double *p_old = parity ? p1.p : p0.p;
  double *p_new = parity ? p0.p : p1.p;
  const int *stop = &sc->done;
  hipEvent_t *ev = profile ? &pev[(size_t)slot * 5] : nullptr;
  if (ev) (void)hipEventRecord(ev[0], stream);
  if (gprec == RCHOL_AMD_FP32) {
    dispatch_sweeps<float, float>(this, r.p, y32.p, z32.p, stop, true, r.p, ev);
    launch_spmv_p<float>(this, z32.p, p_old, p_new);
    if (ev) (void)hipEventRecord(ev[3], stream);
    hipLaunchKernelGGL((k_update<float>), dim3(vec_grid()), dim3(kBlock), 0, stream, (int)n,
                       (const double *)p_new, (const double *)q.p, x.p, r.p, y32.p, partials.p, sc);
  } else {
    dispatch_sweeps<double, double>(this, r.p, y64.p, z64.p, stop, true, r.p, ev);
    launch_spmv_p<double>(this, z64.p, p_old, p_new);
    if (ev) (void)hipEventRecord(ev[3], stream);
    hipLaunchKernelGGL((k_update<double>), dim3(vec_grid()), dim3(kBlock), 0, stream, (int)n,
                       (const double *)p_new, (const double *)q.p, x.p, r.p, y64.p, partials.p,
                       sc);
  }
  if (resid_period > 0) launch_resid_replace(this);
  if (ev) (void)hipEventRecord(ev[4], stream);
  return RCHOL_AMD_OK;
}

int rchol_amd_ctx::pcg(const double *d_b, double tol, int maxit, double *d_x_out, int keep_x,
                       double *relres, int *itr) {
  if (!hasA || !hasG) return fail(RCHOL_AMD_ESTATE, "A and G must be uploaded before pcg");
  if (nG != n) return fail(RCHOL_AMD_EINVAL, "G is %zu x %zu but A is %zu x %zu", nG, nG, n, n);
  int rc = ensure_work();
  if (rc) return rc;
  if (d_x_out != x.p && keep_x)
    HIP_TRY(hipMemcpyAsync(x.p, d_x_out, n * sizeof(double), hipMemcpyDeviceToDevice, stream));
  // scalar set-up
  hsc->tol = tol;
  hsc->maxit = maxit;
  hsc->zero = 0;
  hsc->resid_period = resid_period;
  hsc->strict = strict_sync;
  HIP_TRY(hipMemcpyAsync(&sc->strict, &hsc->strict, sizeof(int), hipMemcpyHostToDevice, stream));
  HIP_TRY(hipMemcpyAsync(&sc->tol, &hsc->tol, sizeof(double), hipMemcpyHostToDevice, stream));
  HIP_TRY(hipMemcpyAsync(&sc->maxit, &hsc->maxit, sizeof(int), hipMemcpyHostToDevice, stream));
  HIP_TRY(hipMemcpyAsync(&sc->zero, &hsc->zero, sizeof(int), hipMemcpyHostToDevice, stream));
  HIP_TRY(hipMemcpyAsync(&sc->resid_period, &hsc->resid_period, sizeof(int),
                         hipMemcpyHostToDevice, stream));
  // the captured replacement kernel reads b from the context's own buffer
  if (resid_period > 0 && d_b != b.p)
    HIP_TRY(hipMemcpyAsync(b.p, d_b, n * sizeof(double), hipMemcpyDeviceToDevice, stream));
  HIP_TRY(hipEventRecord(ev0, stream));
  if (gprec == RCHOL_AMD_FP32)
    hipLaunchKernelGGL((k_init<float>), dim3(vec_grid()), dim3(kBlock), 0, stream, (int)n, d_b,
                       r.p, x.p, keep_x, y32.p, partials.p, sc);
  else
    hipLaunchKernelGGL((k_init<double>), dim3(vec_grid()), dim3(kBlock), 0, stream, (int)n, d_b,
                       r.p, x.p, keep_x, y64.p, partials.p, sc);
  HIP_TRY(hipGetLastError());

  int K = graph_iters > 0 ? graph_iters : 8;
  if (K & 1) K++;
  if (profile && pev.size() < (size_t)K * 5) {
    for (hipEvent_t ev : pev) (void)hipEventDestroy(ev);
    pev.assign((size_t)K * 5, nullptr);
    for (auto &ev : pev) HIP_TRY(hipEventCreate(&ev));
    drop_graph();
  }
  for (int i = 0; i < 4; i++) stats.kernel_ms[i] = 0.0;
  stats.kernel_iters = 0;
  if (graph_iters > 0 && !gexec && !profile) {
    HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < K; i++) launch_iteration(i & 1, i);
    hipError_t e = hipStreamEndCapture(stream, &graph);
    if (e == hipSuccess) e = hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      drop_graph();
      if (!profile) return fail(RCHOL_AMD_EHIP, "graph capture failed: %s", hipGetErrorString(e));
      profile_no_graph = 1;  // event nodes not capturable here: profile with direct launches
    }
  }
  int k_seen = 0;
  for (;;) {
    if (gexec) {
      HIP_TRY(hipGraphLaunch(gexec, stream));
    } else {
      for (int i = 0; i < K; i++) launch_iteration(i & 1, i);
      HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipMemcpyAsync(hsc, sc, sizeof(DevScalars), hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    if (profile) {
      // iterations that ran in this batch occupy slots 0 .. ran-1
      const int ran = std::min(K, hsc->k - k_seen);
      for (int i = 0; i < ran; i++) {
        for (int kk = 0; kk < 4; kk++) {
          float ms = 0.f;
          HIP_TRY(hipEventElapsedTime(&ms, pev[(size_t)i * 5 + kk], pev[(size_t)i * 5 + kk + 1]));
          stats.kernel_ms[kk] += ms;
        }
      }
      stats.kernel_iters += ran;
    }
    k_seen = hsc->k;
    if (hsc->err) break;
    if (hsc->done) break;
  }
  if (hsc->err)
    return stall(hsc->err);
  // the last completed iteration wrote p_new into p1 when k is odd
  launch_spmv_res(this, x.p, d_b, tmp64a.p);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(ev1, stream));
  HIP_TRY(hipMemcpyAsync(hsc, sc, sizeof(DevScalars), hipMemcpyDeviceToHost, stream));
  if (d_x_out != x.p)
    HIP_TRY(hipMemcpyAsync(d_x_out, x.p, n * sizeof(double), hipMemcpyDeviceToDevice, stream));
  HIP_TRY(hipStreamSynchronize(stream));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, ev0, ev1));
  *relres = hsc->relres;
  *itr = hsc->k;
  stats.last_iters = hsc->k;
  stats.last_relres = hsc->relres;
  stats.last_solve_ms = ms;
  return RCHOL_AMD_OK;
}

int rchol_amd_ctx::precond(const double *d_r, double *d_z) {
  if (!hasG) return fail(RCHOL_AMD_ESTATE, "G must be uploaded before precond");
  if (!hasA) n = nG;
  if (n != nG) return fail(RCHOL_AMD_EINVAL, "G is %zu x %zu but A is %zu x %zu", nG, nG, n, n);
  int rc = ensure_work();
  if (rc) return rc;
  hsc->zero = 0;
  hsc->err = 0;
  HIP_TRY(hipMemcpyAsync(&sc->zero, &hsc->zero, sizeof(int), hipMemcpyHostToDevice, stream));
  HIP_TRY(hipMemcpyAsync(&sc->err, &hsc->err, sizeof(int), hipMemcpyHostToDevice, stream));
  const int *stop = &sc->zero;
  const int g = vec_grid();
  if (gprec == RCHOL_AMD_FP32) {
    hipLaunchKernelGGL((k_fill_sentinel<float>), dim3(g), dim3(kBlock), 0, stream, (int)n, y32.p);
    dispatch_sweeps<float, float>(this, d_r, y32.p, z32.p, stop, false, nullptr);
    hipLaunchKernelGGL(k_widen, dim3(g), dim3(kBlock), 0, stream, (int)n, (const float *)z32.p, d_z);
  } else {
    hipLaunchKernelGGL((k_fill_sentinel<double>), dim3(g), dim3(kBlock), 0, stream, (int)n, y64.p);
    dispatch_sweeps<double, double>(this, d_r, y64.p, z64.p, stop, false, nullptr);
    HIP_TRY(hipMemcpyAsync(d_z, z64.p, n * sizeof(double), hipMemcpyDeviceToDevice, stream));
  }
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(&hsc->err, &sc->err, sizeof(int), hipMemcpyDeviceToHost, stream));
  HIP_TRY(hipStreamSynchronize(stream));
  if (hsc->err)
    return stall(hsc->err);
  return RCHOL_AMD_OK;
}


// ---------------------------------------------------------------------------
// Multi-RHS solve (K interleaved columns, per-column recurrences)
// ---------------------------------------------------------------------------
template <typename VT, typename XT, int K>
static int trsv_occupancy_mr() {
  int a = 0, b = 0, c = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
          &a, (const void *)k_sptrsv_mr<VT, XT, double, K, 1>, kBlock, 0) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(
          &b, (const void *)k_sptrsv_mr<VT, XT, XT, K, 2>, kBlock, 0) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(
          &c, (const void *)k_sptrsv_mr<VT, XT, XT, K, 0>, kBlock, 0) != hipSuccess)
    return 0;
  return std::min(a, std::min(b, c));
}

int rchol_amd_ctx::ensure_batch(int K) {
  if (batch.K == K && batch.n == n && batch.b.p && batch.sched_K == K && batch.sched_gen == g_gen)
    return RCHOL_AMD_OK;
  batch.drop_graph();
  const size_t m = n * (size_t)K;
  HIP_TRY(batch.b.alloc(m));
  HIP_TRY(batch.x.alloc(m));
  HIP_TRY(batch.r.alloc(m));
  HIP_TRY(batch.q.alloc(m));
  HIP_TRY(batch.p0.alloc(m));
  HIP_TRY(batch.p1.alloc(m));
  HIP_TRY(batch.res.alloc(m));
  if (gprec == RCHOL_AMD_FP32) {
    batch.y64.reset();
    batch.z64.reset();
    HIP_TRY(batch.y32.alloc(m));
    HIP_TRY(batch.z32.alloc(m));
  } else {
    batch.y32.reset();
    batch.z32.reset();
    HIP_TRY(batch.y64.alloc(m));
    HIP_TRY(batch.z64.alloc(m));
  }
  HIP_TRY(batch.partials.alloc((size_t)ncu * 8 * kMaxCols * 2 + 64));
  if (!batch.bs) {
    HIP_TRY(hipMalloc((void **)&batch.bs, sizeof(BatchScalars)));
    HIP_TRY(hipMemset(batch.bs, 0, sizeof(BatchScalars)));
    HIP_TRY(hipHostMalloc((void **)&batch.hbs, sizeof(BatchScalars), hipHostMallocDefault));
    memset(batch.hbs, 0, sizeof(BatchScalars));
  }
  int per_cu = 0;
  switch (K) {
    case 4: per_cu = gprec == RCHOL_AMD_FP32 ? trsv_occupancy_mr<float, float, 4>() : trsv_occupancy_mr<double, double, 4>(); break;
    case 8: per_cu = gprec == RCHOL_AMD_FP32 ? trsv_occupancy_mr<float, float, 8>() : trsv_occupancy_mr<double, double, 8>(); break;
    case 16: per_cu = gprec == RCHOL_AMD_FP32 ? trsv_occupancy_mr<float, float, 16>() : trsv_occupancy_mr<double, double, 16>(); break;
    default: per_cu = gprec == RCHOL_AMD_FP32 ? trsv_occupancy_mr<float, float, 32>() : trsv_occupancy_mr<double, double, 32>(); break;
  }
  if (per_cu <= 0) return fail(RCHOL_AMD_EHIP, "occupancy query for the batched sweeps failed");
  per_cu = resident_blocks(per_cu);
  batch.K = K;
  batch.n = n;
  return ensure_batch_schedule(K, per_cu);
}

// The K-column sweep schedule for the current G (rebuilt after every
// upload_G or change of K).
int rchol_amd_ctx::ensure_batch_schedule(int K, int per_cu) {
  if (batch.sched_K == K && batch.sched_gen == g_gen) return RCHOL_AMD_OK;
  const int cap_blocks = per_cu * ncu;
  const long long plan_waves = (long long)cap_blocks * (kBlock / 64);
  TaskCut tc[2];
  Sweep *sw[2] = {&fwd, &bwd};
  size_t max_tasks = 0;
  for (int k = 0; k < 2; k++) {
    cut_tasks_mr(sw[k]->level_ptr_h, sw[k]->rp_h, sw[k]->nlevels, K, plan_waves, tc[k]);
    max_tasks = std::max(max_tasks, tc[k].s0.size());
  }
  batch.grid = (int)std::max<size_t>(1, std::min<size_t>(cap_blocks, (max_tasks + 3) / 4));
  HIP_TRY(hipMemset(&batch.bs->arrive, 0, sizeof(batch.bs->arrive)));
  for (int k = 0; k < 2; k++) {
    std::vector<int> wp, ws0, winfo;
    assign_waves(tc[k], batch.grid, wp, ws0, winfo);
    HIP_TRY(batch.wave_ptr[k].upload(wp.data(), wp.size()));
    HIP_TRY(batch.wt_s0[k].upload(ws0.data(), ws0.size()));
    HIP_TRY(batch.wt_info[k].upload(winfo.data(), winfo.size()));
    batch.ntasks[k] = (int)tc[k].s0.size();
  }
  batch.sched_K = K;
  batch.sched_gen = g_gen;
  return RCHOL_AMD_OK;
}

template <int K>
void rchol_amd_ctx::launch_sweeps_mr(const double *rhs, bool with_dot, const int *stop,
                                     hipEvent_t *ev) {
  const dim3 g(batch.grid), blk(kBlock);
  const int *wpf = batch.wave_ptr[0].p, *wsf = batch.wt_s0[0].p, *wif = batch.wt_info[0].p;
  const int *wpb = batch.wave_ptr[1].p, *wsb = batch.wt_s0[1].p, *wib = batch.wt_info[1].p;
  if (gprec == RCHOL_AMD_FP32) {
    hipLaunchKernelGGL((k_sptrsv_mr<float, float, double, K, 1>), g, blk, 0, stream, wpf, wsf,
                       wif, fwd.slot_row.p, fwd.rp.p, fwd.ci.p, fwd.cv32.p, fwd.diag32.p, rhs,
                       batch.y32.p, batch.z32.p, nullptr, nullptr, batch.bs, stop);
    if (ev) (void)hipEventRecord(ev[1], stream);
    if (with_dot)
      hipLaunchKernelGGL((k_sptrsv_mr<float, float, float, K, 2>), g, blk, 0, stream, wpb, wsb,
                         wib, bwd.slot_row.p, bwd.rp.p, bwd.ci.p, bwd.cv32.p, bwd.diag32.p,
                         (const float *)batch.y32.p, batch.z32.p, (float *)nullptr,
                         (const double *)batch.r.p, batch.partials.p, batch.bs, stop);
    else
      hipLaunchKernelGGL((k_sptrsv_mr<float, float, float, K, 0>), g, blk, 0, stream, wpb, wsb,
                         wib, bwd.slot_row.p, bwd.rp.p, bwd.ci.p, bwd.cv32.p, bwd.diag32.p,
                         (const float *)batch.y32.p, batch.z32.p, (float *)nullptr, nullptr,
                         nullptr, batch.bs, stop);
  } else {
    hipLaunchKernelGGL((k_sptrsv_mr<double, double, double, K, 1>), g, blk, 0, stream, wpf, wsf,
                       wif, fwd.slot_row.p, fwd.rp.p, fwd.ci.p, fwd.cv64.p, fwd.diag64.p, rhs,
                       batch.y64.p, batch.z64.p, nullptr, nullptr, batch.bs, stop);
    if (ev) (void)hipEventRecord(ev[1], stream);
    if (with_dot)
      hipLaunchKernelGGL((k_sptrsv_mr<double, double, double, K, 2>), g, blk, 0, stream, wpb,
                         wsb, wib, bwd.slot_row.p, bwd.rp.p, bwd.ci.p, bwd.cv64.p, bwd.diag64.p,
                         (const double *)batch.y64.p, batch.z64.p, (double *)nullptr,
                         (const double *)batch.r.p, batch.partials.p, batch.bs, stop);
    else
      hipLaunchKernelGGL((k_sptrsv_mr<double, double, double, K, 0>), g, blk, 0, stream, wpb,
                         wsb, wib, bwd.slot_row.p, bwd.rp.p, bwd.ci.p, bwd.cv64.p, bwd.diag64.p,
                         (const double *)batch.y64.p, batch.z64.p, (double *)nullptr, nullptr,
                         nullptr, batch.bs, stop);
  }
  if (ev) (void)hipEventRecord(ev[2], stream);
}

template <int K>
void rchol_amd_ctx::launch_iteration_mr(int parity, int slot) {
  double *p_old = parity ? batch.p1.p : batch.p0.p;
  double *p_new = parity ? batch.p0.p : batch.p1.p;
  hipEvent_t *ev = profile ? &pev[(size_t)slot * 5] : nullptr;
  if (ev) (void)hipEventRecord(ev[0], stream);
  launch_sweeps_mr<K>(batch.r.p, true, &batch.bs->all_done, ev);
  const int rows_per_block = (kBlock / 64) * (64 / K);
  int sg = (int)std::max<size_t>(
      1, std::min<size_t>((n + rows_per_block - 1) / rows_per_block, (size_t)ncu * kSpmvMrBlocksPerCu));
  if (sg > 8) sg = (sg + 7) / 8 * 8;  // whole XCD rounds for the row walk
  const int vg = (int)std::max<size_t>(
      1, std::min<size_t>((n * K + kBlock - 1) / kBlock, (size_t)ncu * kVecBlocksPerCu));
  if (gprec == RCHOL_AMD_FP32) {
    hipLaunchKernelGGL((k_spmv_p_mr<float, K>), dim3(sg), dim3(kBlock), 0, stream, (int)n,
                       arp.p, aci.p, av.p, (const float *)batch.z32.p, (const double *)p_old,
                       p_new, batch.q.p, (const double *)batch.r.p, batch.partials.p, batch.bs);
    if (ev) (void)hipEventRecord(ev[3], stream);
    hipLaunchKernelGGL((k_update_mr<float, K>), dim3(vg), dim3(kBlock), 0, stream, (int)n,
                       (const double *)p_new, (const double *)batch.q.p, batch.x.p, batch.r.p,
                       batch.y32.p, batch.partials.p, batch.bs);
  } else {
    hipLaunchKernelGGL((k_spmv_p_mr<double, K>), dim3(sg), dim3(kBlock), 0, stream, (int)n,
                       arp.p, aci.p, av.p, (const double *)batch.z64.p, (const double *)p_old,
                       p_new, batch.q.p, (const double *)batch.r.p, batch.partials.p, batch.bs);
    if (ev) (void)hipEventRecord(ev[3], stream);
    hipLaunchKernelGGL((k_update_mr<double, K>), dim3(vg), dim3(kBlock), 0, stream, (int)n,
                       (const double *)p_new, (const double *)batch.q.p, batch.x.p, batch.r.p,
                       batch.y64.p, batch.partials.p, batch.bs);
  }
  if (ev) (void)hipEventRecord(ev[4], stream);
}

template <int K>
int rchol_amd_ctx::pcg_batch(const double *d_b, int ncols, double tol, int maxit, double *d_x,
                             int keep_x, double *relres, int *itr) {
  if (!hasA || !hasG) return fail(RCHOL_AMD_ESTATE, "A and G must be uploaded before pcg");
  if (nG != n) return fail(RCHOL_AMD_EINVAL, "G is %zu x %zu but A is %zu x %zu", nG, nG, n, n);
  int rc = ensure_batch(K);
  if (rc) return rc;
  const size_t m = n * (size_t)K;
  if (keep_x && d_x != batch.x.p)
    HIP_TRY(hipMemcpyAsync(batch.x.p, d_x, m * sizeof(double), hipMemcpyDeviceToDevice, stream));
  batch.hbs->tol = tol;
  batch.hbs->maxit = maxit;
  batch.hbs->ncols = ncols;
  batch.hbs->zero = 0;
  batch.hbs->strict = strict_sync;
  HIP_TRY(hipMemcpyAsync(&batch.bs->strict, &batch.hbs->strict, sizeof(int), hipMemcpyHostToDevice, stream));
  HIP_TRY(hipMemcpyAsync(&batch.bs->tol, &batch.hbs->tol, sizeof(double), hipMemcpyHostToDevice, stream));
  HIP_TRY(hipMemcpyAsync(&batch.bs->maxit, &batch.hbs->maxit, sizeof(int), hipMemcpyHostToDevice, stream));
  HIP_TRY(hipMemcpyAsync(&batch.bs->ncols, &batch.hbs->ncols, sizeof(int), hipMemcpyHostToDevice, stream));
  HIP_TRY(hipMemcpyAsync(&batch.bs->zero, &batch.hbs->zero, sizeof(int), hipMemcpyHostToDevice, stream));
  HIP_TRY(hipEventRecord(ev0, stream));
  const int vg = (int)std::max<size_t>(1, std::min<size_t>((m + kBlock - 1) / kBlock, (size_t)ncu * kVecBlocksPerCu));
  if (gprec == RCHOL_AMD_FP32)
    hipLaunchKernelGGL((k_init_mr<float, K>), dim3(vg), dim3(kBlock), 0, stream, (int)n, d_b,
                       batch.r.p, batch.x.p, keep_x, batch.y32.p, batch.partials.p, batch.bs);
  else
    hipLaunchKernelGGL((k_init_mr<double, K>), dim3(vg), dim3(kBlock), 0, stream, (int)n, d_b,
                       batch.r.p, batch.x.p, keep_x, batch.y64.p, batch.partials.p, batch.bs);
  HIP_TRY(hipGetLastError());
  int G = graph_iters > 0 ? graph_iters : 8;
  if (G & 1) G++;
  if (profile && pev.size() < (size_t)G * 5) {
    for (hipEvent_t ev : pev) (void)hipEventDestroy(ev);
    pev.assign((size_t)G * 5, nullptr);
    for (auto &ev : pev) HIP_TRY(hipEventCreate(&ev));
    drop_graph();
  }
  for (int i = 0; i < 4; i++) stats.kernel_ms[i] = 0.0;
  stats.kernel_iters = 0;
  if (graph_iters > 0 && !batch.gexec && !profile) {
    HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < G; i++) launch_iteration_mr<K>(i & 1, i);
    hipError_t e = hipStreamEndCapture(stream, &batch.graph);
    if (e == hipSuccess) e = hipGraphInstantiate(&batch.gexec, batch.graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      batch.drop_graph();
      if (!profile) return fail(RCHOL_AMD_EHIP, "graph capture failed: %s", hipGetErrorString(e));
      profile_no_graph = 1;
    }
  }
  int maxk_seen = 0;
  for (;;) {
    if (batch.gexec) {
      HIP_TRY(hipGraphLaunch(batch.gexec, stream));
    } else {
      for (int i = 0; i < G; i++) launch_iteration_mr<K>(i & 1, i);
      HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipMemcpyAsync(batch.hbs, batch.bs, sizeof(BatchScalars), hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    int maxk = 0;
    for (int c = 0; c < ncols; c++) maxk = std::max(maxk, batch.hbs->k[c]);
    if (profile) {
      const int ran = std::min(G, maxk - maxk_seen);
      for (int i = 0; i < ran; i++)
        for (int kk = 0; kk < 4; kk++) {
          float ms = 0.f;
          HIP_TRY(hipEventElapsedTime(&ms, pev[(size_t)i * 5 + kk], pev[(size_t)i * 5 + kk + 1]));
          stats.kernel_ms[kk] += ms;
        }
      stats.kernel_iters += ran;
    }
    maxk_seen = maxk;
    if (batch.hbs->err || batch.hbs->all_done) break;
  }
  if (batch.hbs->err)
    return stall(batch.hbs->err);
  const int rows_per_block = (kBlock / 64) * (64 / K);
  int sg = (int)std::max<size_t>(
      1, std::min<size_t>((n + rows_per_block - 1) / rows_per_block, (size_t)ncu * kSpmvMrBlocksPerCu));
  if (sg > 8) sg = (sg + 7) / 8 * 8;  // whole XCD rounds for the row walk
  hipLaunchKernelGGL((k_spmv_res_mr<K>), dim3(sg), dim3(kBlock), 0, stream, (int)n, arp.p, aci.p,
                     av.p, (const double *)batch.x.p, d_b, batch.res.p, batch.partials.p, batch.bs);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(ev1, stream));
  HIP_TRY(hipMemcpyAsync(batch.hbs, batch.bs, sizeof(BatchScalars), hipMemcpyDeviceToHost, stream));
  if (d_x != batch.x.p)
    HIP_TRY(hipMemcpyAsync(d_x, batch.x.p, m * sizeof(double), hipMemcpyDeviceToDevice, stream));
  HIP_TRY(hipStreamSynchronize(stream));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, ev0, ev1));
  int maxk = 0;
  for (int c = 0; c < ncols; c++) {
    relres[c] = batch.hbs->relres[c];
    itr[c] = batch.hbs->k[c];
    maxk = std::max(maxk, itr[c]);
  }
  stats.last_iters = maxk;
  stats.last_relres = ncols ? relres[0] : 0.0;
  stats.last_solve_ms = ms;
  return RCHOL_AMD_OK;
}

template <int K>
int rchol_amd_ctx::precond_batch(const double *d_r, double *d_z) {
  if (!hasG) return fail(RCHOL_AMD_ESTATE, "G must be uploaded before precond");
  if (!hasA) n = nG;
  if (n != nG) return fail(RCHOL_AMD_EINVAL, "G is %zu x %zu but A is %zu x %zu", nG, nG, n, n);
  int rc = ensure_batch(K);
  if (rc) return rc;
  const size_t m = n * (size_t)K;
  batch.hbs->zero = 0;
  batch.hbs->err = 0;
  HIP_TRY(hipMemcpyAsync(&batch.bs->zero, &batch.hbs->zero, sizeof(int), hipMemcpyHostToDevice, stream));
  HIP_TRY(hipMemcpyAsync(&batch.bs->err, &batch.hbs->err, sizeof(int), hipMemcpyHostToDevice, stream));
  const int g = (int)std::max<size_t>(1, std::min<size_t>((m + kBlock - 1) / kBlock, (size_t)ncu * 8));
  if (gprec == RCHOL_AMD_FP32) {
    hipLaunchKernelGGL((k_fill_sentinel<float>), dim3(g), dim3(kBlock), 0, stream, (int)m, batch.y32.p);
    launch_sweeps_mr<K>(d_r, false, &batch.bs->zero, nullptr);
    hipLaunchKernelGGL(k_widen, dim3(g), dim3(kBlock), 0, stream, (int)m, (const float *)batch.z32.p, d_z);
  } else {
    hipLaunchKernelGGL((k_fill_sentinel<double>), dim3(g), dim3(kBlock), 0, stream, (int)m, batch.y64.p);
    launch_sweeps_mr<K>(d_r, false, &batch.bs->zero, nullptr);
    HIP_TRY(hipMemcpyAsync(d_z, batch.z64.p, m * sizeof(double), hipMemcpyDeviceToDevice, stream));
  }
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(&batch.hbs->err, &batch.bs->err, sizeof(int), hipMemcpyDeviceToHost, stream));
  HIP_TRY(hipStreamSynchronize(stream));
  if (batch.hbs->err)
    return stall(batch.hbs->err);
  return RCHOL_AMD_OK;
}

static int batch_width(int k) {
  if (k <= 4) return 4;
  if (k <= 8) return 8;
  if (k <= 16) return 16;
  return 32;
}

// Dispatch on the compile-time width.
#define RCHOL_BATCH_DISPATCH(K, CALL)                  \
  switch (K) {                                          \
    case 4: { constexpr int KK = 4; return c->CALL; }   \
    case 8: { constexpr int KK = 8; return c->CALL; }   \
    case 16: { constexpr int KK = 16; return c->CALL; } \
    default: { constexpr int KK = 32; return c->CALL; } \
  }

static int pcg_batch_dev(rchol_amd_ctx *c, int K, const double *d_b, int k, double tol, int maxit,
                         double *d_x, int keep, double *relres, int *itr) {
  RCHOL_BATCH_DISPATCH(K, template pcg_batch<KK>(d_b, k, tol, maxit, d_x, keep, relres, itr));
}

static int precond_batch_dev(rchol_amd_ctx *c, int K, const double *d_r, double *d_z) {
  RCHOL_BATCH_DISPATCH(K, template precond_batch<KK>(d_r, d_z));
}


// ---------------------------------------------------------------------------
// Uneven batch: stacked independent systems with per-system scalars
// ---------------------------------------------------------------------------
namespace {
template <int L>
void launch_spmv_p_seg_t(rchol_amd_ctx *c, SegLayout lay, const void *z, bool z32,
                         const double *p_old, double *p_new) {
  if (L == 0) {  // stream form: p update, then the SpMV gathering p'
    if (z32)
      hipLaunchKernelGGL((k_p_update_seg<float>), dim3(c->segs.nblocks), dim3(kBlock), 0,
                         c->stream, lay, (const float *)z, p_old, p_new, c->segs.ss);
    else
      hipLaunchKernelGGL((k_p_update_seg<double>), dim3(c->segs.nblocks), dim3(kBlock), 0,
                         c->stream, lay, (const double *)z, p_old, p_new, c->segs.ss);
    hipLaunchKernelGGL(k_spmv_p_seg_s, dim3(c->segs.nblocks), dim3(kBlock), 0, c->stream, lay,
                       c->arp.p, c->aci.p, c->av.p, (const double *)p_new, c->q.p,
                       (const double *)c->r.p, c->partials.p, c->segs.ss);
    return;
  }
  constexpr int LL = L ? L : 8;
  if (z32)
    hipLaunchKernelGGL((k_spmv_p_seg<float, LL>), dim3(c->segs.nblocks), dim3(kBlock), 0,
                       c->stream, lay, c->arp.p, c->aci.p, c->av.p, (const float *)z, p_old, p_new,
                       c->q.p, (const double *)c->r.p, c->partials.p, c->segs.ss);
  else
    hipLaunchKernelGGL((k_spmv_p_seg<double, LL>), dim3(c->segs.nblocks), dim3(kBlock), 0,
                       c->stream, lay, c->arp.p, c->aci.p, c->av.p, (const double *)z, p_old,
                       p_new, c->q.p, (const double *)c->r.p, c->partials.p, c->segs.ss);
}
SegLayout seg_layout(rchol_amd_ctx *c) {
  return SegLayout{c->segs.seg.p, c->segs.r0.p, c->segs.r1.p, c->segs.sb.p};
}
}  // namespace

void rchol_amd_ctx::launch_iteration_seg(int parity, int slot) {
  double *p_old = parity ? p1.p : p0.p;
  double *p_new = parity ? p0.p : p1.p;
  const int *stop = &segs.ss->all_done;
  hipEvent_t *ev = profile ? &pev[(size_t)slot * 5] : nullptr;
  const SegLayout lay = seg_layout(this);
  const bool f32 = gprec == RCHOL_AMD_FP32;
  if (ev) (void)hipEventRecord(ev[0], stream);
  if (f32)
    dispatch_sweeps<float, float>(this, r.p, y32.p, z32.p, stop, false, nullptr, ev);
  else
    dispatch_sweeps<double, double>(this, r.p, y64.p, z64.p, stop, false, nullptr, ev);
  const dim3 g(segs.nblocks), blk(kBlock);
  if (f32)
    hipLaunchKernelGGL((k_dot_seg<float>), g, blk, 0, stream, lay, (const double *)r.p,
                       (const float *)z32.p, partials.p, segs.ss);
  else
    hipLaunchKernelGGL((k_dot_seg<double>), g, blk, 0, stream, lay, (const double *)r.p,
                       (const double *)z64.p, partials.p, segs.ss);
  const void *z = f32 ? (const void *)z32.p : (const void *)z64.p;
  switch (spmv_lanes) {
    case 0: launch_spmv_p_seg_t<0>(this, lay, z, f32, p_old, p_new); break;
    case 4: launch_spmv_p_seg_t<4>(this, lay, z, f32, p_old, p_new); break;
    case 16: launch_spmv_p_seg_t<16>(this, lay, z, f32, p_old, p_new); break;
    case 32: launch_spmv_p_seg_t<32>(this, lay, z, f32, p_old, p_new); break;
    default: launch_spmv_p_seg_t<8>(this, lay, z, f32, p_old, p_new); break;
  }
  if (ev) (void)hipEventRecord(ev[3], stream);
  if (f32)
    hipLaunchKernelGGL((k_update_seg<float>), g, blk, 0, stream, lay, (const double *)p_new,
                       (const double *)q.p, x.p, r.p, y32.p, partials.p, segs.ss);
  else
    hipLaunchKernelGGL((k_update_seg<double>), g, blk, 0, stream, lay, (const double *)p_new,
                       (const double *)q.p, x.p, r.p, y64.p, partials.p, segs.ss);
  if (ev) (void)hipEventRecord(ev[4], stream);
}

int rchol_amd_ctx::pcg_segments(const double *d_b, double tol, int maxit, double *d_x, int keep_x,
                                double *relres, int *itr) {
  if (!hasA || !hasG) return fail(RCHOL_AMD_ESTATE, "A and G must be uploaded before pcg");
  if (segs.S == 0 || segs.off.back() != n)
    return fail(RCHOL_AMD_ESTATE, "segments must be set for the uploaded A");
  int rc = ensure_work();
  if (rc) return rc;
  if (keep_x && d_x != x.p)
    HIP_TRY(hipMemcpyAsync(x.p, d_x, n * sizeof(double), hipMemcpyDeviceToDevice, stream));
  segs.hss->tol = tol;
  segs.hss->maxit = maxit;
  segs.hss->nseg = segs.S;
  segs.hss->strict = strict_sync;
  HIP_TRY(hipMemcpyAsync(&segs.ss->strict, &segs.hss->strict, sizeof(int), hipMemcpyHostToDevice, stream));
  HIP_TRY(hipMemcpyAsync(&segs.ss->tol, &segs.hss->tol, sizeof(double), hipMemcpyHostToDevice, stream));
  HIP_TRY(hipMemcpyAsync(&segs.ss->maxit, &segs.hss->maxit, sizeof(int), hipMemcpyHostToDevice, stream));
  HIP_TRY(hipMemcpyAsync(&segs.ss->nseg, &segs.hss->nseg, sizeof(int), hipMemcpyHostToDevice, stream));
  hsc->err = 0;
  HIP_TRY(hipMemcpyAsync(&sc->err, &hsc->err, sizeof(int), hipMemcpyHostToDevice, stream));
  HIP_TRY(hipEventRecord(ev0, stream));
  const SegLayout lay = seg_layout(this);
  if (gprec == RCHOL_AMD_FP32)
    hipLaunchKernelGGL((k_init_seg<float>), dim3(segs.nblocks), dim3(kBlock), 0, stream, lay, d_b,
                       r.p, x.p, keep_x, y32.p, partials.p, segs.ss);
  else
    hipLaunchKernelGGL((k_init_seg<double>), dim3(segs.nblocks), dim3(kBlock), 0, stream, lay,
                       d_b, r.p, x.p, keep_x, y64.p, partials.p, segs.ss);
  HIP_TRY(hipGetLastError());
  int G = graph_iters > 0 ? graph_iters : 8;
  if (G & 1) G++;
  if (profile && pev.size() < (size_t)G * 5) {
    for (hipEvent_t ev : pev) (void)hipEventDestroy(ev);
    pev.assign((size_t)G * 5, nullptr);
    for (auto &ev : pev) HIP_TRY(hipEventCreate(&ev));
    drop_graph();
  }
  for (int i = 0; i < 4; i++) stats.kernel_ms[i] = 0.0;
  stats.kernel_iters = 0;
  if (graph_iters > 0 && !segs.gexec && !profile) {
    HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < G; i++) launch_iteration_seg(i & 1, i);
    hipError_t e = hipStreamEndCapture(stream, &segs.graph);
    if (e == hipSuccess) e = hipGraphInstantiate(&segs.gexec, segs.graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      segs.drop_graph();
      if (!profile) return fail(RCHOL_AMD_EHIP, "graph capture failed: %s", hipGetErrorString(e));
      profile_no_graph = 1;
    }
  }
  int maxk_seen = 0;
  for (;;) {
    if (segs.gexec) {
      HIP_TRY(hipGraphLaunch(segs.gexec, stream));
    } else {
      for (int i = 0; i < G; i++) launch_iteration_seg(i & 1, i);
      HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipMemcpyAsync(segs.hss, segs.ss, sizeof(SegScalars), hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipMemcpyAsync(&hsc->err, &sc->err, sizeof(int), hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    int maxk = 0;
    for (int s = 0; s < segs.S; s++) maxk = std::max(maxk, segs.hss->k[s]);
    if (profile) {
      const int ran = std::min(G, maxk - maxk_seen);
      for (int i = 0; i < ran; i++)
        for (int kk = 0; kk < 4; kk++) {
          float ms = 0.f;
          HIP_TRY(hipEventElapsedTime(&ms, pev[(size_t)i * 5 + kk], pev[(size_t)i * 5 + kk + 1]));
          stats.kernel_ms[kk] += ms;
        }
      stats.kernel_iters += ran;
    }
    maxk_seen = maxk;
    if (hsc->err || segs.hss->all_done) break;
  }
  if (hsc->err)
    return stall(hsc->err);
  switch (spmv_lanes) {
    case 0: hipLaunchKernelGGL(k_spmv_res_seg_s, dim3(segs.nblocks), dim3(kBlock), 0, stream, lay, arp.p, aci.p, av.p, (const double *)x.p, d_b, tmp64a.p, partials.p, segs.ss); break;
    case 4: hipLaunchKernelGGL((k_spmv_res_seg<4>), dim3(segs.nblocks), dim3(kBlock), 0, stream, lay, arp.p, aci.p, av.p, (const double *)x.p, d_b, tmp64a.p, partials.p, segs.ss); break;
    case 16: hipLaunchKernelGGL((k_spmv_res_seg<16>), dim3(segs.nblocks), dim3(kBlock), 0, stream, lay, arp.p, aci.p, av.p, (const double *)x.p, d_b, tmp64a.p, partials.p, segs.ss); break;
    case 32: hipLaunchKernelGGL((k_spmv_res_seg<32>), dim3(segs.nblocks), dim3(kBlock), 0, stream, lay, arp.p, aci.p, av.p, (const double *)x.p, d_b, tmp64a.p, partials.p, segs.ss); break;
    default: hipLaunchKernelGGL((k_spmv_res_seg<8>), dim3(segs.nblocks), dim3(kBlock), 0, stream, lay, arp.p, aci.p, av.p, (const double *)x.p, d_b, tmp64a.p, partials.p, segs.ss); break;
  }
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(ev1, stream));
  HIP_TRY(hipMemcpyAsync(segs.hss, segs.ss, sizeof(SegScalars), hipMemcpyDeviceToHost, stream));
  if (d_x != x.p)
    HIP_TRY(hipMemcpyAsync(d_x, x.p, n * sizeof(double), hipMemcpyDeviceToDevice, stream));
  HIP_TRY(hipStreamSynchronize(stream));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, ev0, ev1));
  int maxk = 0;
  for (int s = 0; s < segs.S; s++) {
    relres[s] = segs.hss->relres[s];
    itr[s] = segs.hss->k[s];
    maxk = std::max(maxk, itr[s]);
  }
  stats.last_iters = maxk;
  stats.last_relres = relres[0];
  stats.last_solve_ms = ms;
  return RCHOL_AMD_OK;
}

namespace {

// scratch's address on each device (null until looked up there)
std::atomic<unsigned char *> g_scratch_dev[kMaxDevices];
std::mutex g_scratch_mu;

// Look up scratch's address on `device`, which must be current; once per
// device.
int scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices)
    return fail(RCHOL_AMD_EINVAL, "device %d beyond the %d the engine tracks", device, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_dev[device].load()) return RCHOL_AMD_OK;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(RCHOL_AMD_EHIP, "hipGetSymbolAddress(scratch) on device %d failed: %s", device,
                hipGetErrorString(e));
  g_scratch_dev[device].store((unsigned char *)p);
  return RCHOL_AMD_OK;
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Byte offsets inside scratch for an n-row factor: [control block | y | z
// (FP32 factor only; FP64 sweeps write z into the caller's output) | r(P),
// z(P) staging (when a permutation is set)].  Returns the bytes needed.
struct ScratchLayout {
  size_t y = 0, z = 0, pin = 0, pout = 0, bytes = 0;
};
ScratchLayout scratch_layout(size_t n, bool fp32, bool perm) {
  ScratchLayout L;
  size_t o = kScratchCtl;
  L.y = o;
  o += align256(n * (fp32 ? sizeof(float) : sizeof(double)));
  if (fp32) {
    L.z = o;
    o += align256(n * sizeof(float));
  }
  if (perm) {
    L.pin = o;
    o += align256(n * sizeof(double));
    L.pout = o;
    o += align256(n * sizeof(double));
  }
  L.bytes = o;
  return L;
}

// Zero the control block (a kernel, not a memset, so a captured graph can
// hold it): no error, stop flag 0, counters 0.
__global__ void k_clear_ctl(unsigned int *ctl, int words) {
  for (int i = threadIdx.x; i < words; i += blockDim.x) ctl[i] = 0u;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char *rchol_amd_last_error(void) { return g_err.c_str(); }

int rchol_amd_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int rchol_amd_create(int device, rchol_amd_ctx **out) {
  if (!out) return fail(RCHOL_AMD_EINVAL, "null output pointer");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(RCHOL_AMD_EHIP, "no HIP device is visible");
  if (device < 0 || device >= ndev)
    return fail(RCHOL_AMD_EINVAL, "device %d out of range (%d visible)", device, ndev);
  std::unique_ptr<rchol_amd_ctx> c(new rchol_amd_ctx());
  c->device = device;
  HIP_TRY(hipSetDevice(device));
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, device));
  c->ncu = prop.multiProcessorCount;
  c->lds_max = (int)prop.sharedMemPerBlock;
  HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  HIP_TRY(hipMalloc((void **)&c->sc, sizeof(DevScalars)));
  HIP_TRY(hipMemset(c->sc, 0, sizeof(DevScalars)));
  HIP_TRY(hipHostMalloc((void **)&c->hsc, sizeof(DevScalars), hipHostMallocDefault));
  memset(c->hsc, 0, sizeof(DevScalars));
  HIP_TRY(hipEventCreate(&c->ev0));
  HIP_TRY(hipEventCreate(&c->ev1));
  const int rc = scratch_init(device);
  if (rc) return rc;
  *out = c.release();
  return RCHOL_AMD_OK;
}

int rchol_amd_destroy(rchol_amd_ctx *ctx) {
  if (!ctx) return RCHOL_AMD_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  delete ctx;
  return RCHOL_AMD_OK;
}

int rchol_amd_upload_A(rchol_amd_ctx *c, size_t n, const size_t *rowPtr, const size_t *colIdx,
                       const double *val) {
  if (!c || !rowPtr || (!colIdx && n) || (!val && n)) return fail(RCHOL_AMD_EINVAL, "null argument");
  if (n == 0 || n > kIdxMax) return fail(RCHOL_AMD_ERANGE, "n=%zu outside [1, 2^31)", n);
  const size_t nnz = rowPtr[n];
  if (rowPtr[0] != 0) return fail(RCHOL_AMD_EINVAL, "rowPtr[0] must be 0");
  if (nnz > kIdxMax) return fail(RCHOL_AMD_ERANGE, "nnz(A)=%zu exceeds the int32 index range", nnz);
  const double t0 = host_ms();
  std::vector<int> rp(n + 1);
  std::vector<int, NoInit<int>> ci(nnz);
  for (size_t i = 0; i <= n; i++) {
    if (i && rowPtr[i] < rowPtr[i - 1]) return fail(RCHOL_AMD_EINVAL, "rowPtr not monotone at %zu", i);
    rp[i] = (int)rowPtr[i];
  }
  size_t bad = nnz;  // lowest out-of-range column index
  par_for(nnz, [&](size_t e0, size_t e1) {
    for (size_t e = e0; e < e1; e++) {
      if (colIdx[e] >= n) {
        size_t cur = __atomic_load_n(&bad, __ATOMIC_RELAXED);
        while (e < cur &&
               !__atomic_compare_exchange_n(&bad, &cur, e, false, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
        }
        return;
      }
      ci[e] = (int)colIdx[e];
    }
  });
  if (bad < nnz) return fail(RCHOL_AMD_EINVAL, "colIdx[%zu]=%zu out of range", bad, colIdx[bad]);
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->drop_graph();
  c->hasA = false;
  HIP_TRY(c->arp.upload(rp.data(), n + 1));
  HIP_TRY(c->aci.upload(ci.data(), nnz));
  HIP_TRY(c->av.upload(val, nnz));
  c->n = n;
  c->nnzA = nnz;
  c->hasA = true;
  c->stats.n = n;
  c->stats.nnz_a = nnz;
  c->stats.setup_ms[0] = host_ms() - t0;
  return RCHOL_AMD_OK;
}

int rchol_amd_update_A_values(rchol_amd_ctx *c, size_t nnz, const double *val) {
  if (!c || (!val && nnz)) return fail(RCHOL_AMD_EINVAL, "null argument");
  if (!c->hasA) return fail(RCHOL_AMD_ESTATE, "no matrix is resident");
  if (nnz != c->nnzA)
    return fail(RCHOL_AMD_EINVAL, "nnz=%zu differs from the resident matrix's %zu", nnz, c->nnzA);
  const double t0 = host_ms();
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipStreamSynchronize(c->stream));
  // the values change in place: captured graphs keep their pointers
  HIP_TRY(hipMemcpy(c->av.p, val, nnz * sizeof(double), hipMemcpyHostToDevice));
  c->stats.setup_ms[0] = host_ms() - t0;
  return RCHOL_AMD_OK;
}

}  // extern "C"

namespace {

// A factor is "deep" when its levels are many and narrow on average: then
// the chunked sweeps are planned and their estimate compared with the level
// schedule's (about kLevelUs per level, measured on MI355X).
constexpr int kChunkMinLevels = 1024;
// Rows per chunk: at most what the LDS holds and at most kChunkMaxRows (the
// local row index has 16 bits).  20352 is what 160 KB of LDS holds in FP64;
// FP32 chunks of that size run faster than LDS-sized ones (256^3 Poisson in
// the reference's order: 37.6 vs 40.7 ms per forward sweep).
#ifndef RCHOL_AMD_CHUNK_ROWS
#define RCHOL_AMD_CHUNK_ROWS 20352
#endif
constexpr size_t kChunkMaxRows = RCHOL_AMD_CHUNK_ROWS;
static_assert(kChunkMaxRows <= 65535, "chunk rows index 16-bit local rows");
constexpr double kChunkMaxMeanWidth = 4096.0;
constexpr double kLevelUs = 1.3;

template <typename T, typename Vec>
hipError_t upload_as(DevBuf<T> &d, const Vec &h) {
  if constexpr (sizeof(T) == sizeof(double)) {
    return d.upload((const T *)h.data(), h.size());
  } else {
    std::vector<T, NoInit<T>> f(h.size());
    par_for(h.size(), [&](size_t e0, size_t e1) {
      for (size_t e = e0; e < e1; e++) f[e] = (T)h[e];
    });
    return d.upload(f.data(), f.size());
  }
}

int upload_chunks(Sweep::Chunks &d, const HostChunks &h, int precision) {
  d.reset();
  HIP_TRY(d.order.upload(h.order.data(), h.order.size()));
  HIP_TRY(d.q0.upload(h.q0.data(), h.q0.size()));
  HIP_TRY(d.m.upload(h.m.data(), h.m.size()));
  HIP_TRY(d.r0.upload(h.r0.data(), h.r0.size()));
  HIP_TRY(d.r1.upload(h.r1.data(), h.r1.size()));
  HIP_TRY(d.split.upload(h.split.data(), h.split.size()));
  HIP_TRY(d.q_row.upload(h.q_row.data(), h.q_row.size()));
  HIP_TRY(d.x_l.upload(h.x_l.data(), h.x_l.size()));
  HIP_TRY(d.x_ci.upload(h.x_ci.data(), h.x_ci.size()));
  HIP_TRY(d.r_w.upload(h.r_w.data(), h.r_w.size()));
  HIP_TRY(d.dep_ptr.upload(h.dep_ptr.data(), h.dep_ptr.size()));
  HIP_TRY(d.dep.upload(h.dep.data(), h.dep.size()));
  const int items = h.nchunks + h.nparts;
  HIP_TRY(d.done.alloc(items));
  HIP_TRY(hipMemset(d.done.p, 0, std::max(items, 1) * sizeof(int)));
  HIP_TRY(d.c_part0.upload(h.c_part0.data(), h.c_part0.size()));
  HIP_TRY(d.p_chunk.upload(h.p_chunk.data(), h.p_chunk.size()));
  HIP_TRY(d.p_l0.upload(h.p_l0.data(), h.p_l0.size()));
  HIP_TRY(d.p_l1.upload(h.p_l1.data(), h.p_l1.size()));
  HIP_TRY(d.p_lg.upload(h.p_lg.data(), h.p_lg.size()));
  HIP_TRY(d.x_rp.upload(h.x_rp.data(), h.x_rp.size()));
  if (precision == RCHOL_AMD_FP32)
    HIP_TRY(d.ext32.alloc(h.q_row.size()));
  else
    HIP_TRY(d.ext64.alloc(h.q_row.size()));
  d.nparts = h.nparts;
  if (precision == RCHOL_AMD_FP32) {
    HIP_TRY(upload_as(d.q_diag32, h.q_diag));
    HIP_TRY(upload_as(d.x_cv32, h.x_cv));
    HIP_TRY(upload_as(d.r_v32, h.r_v));
  } else {
    HIP_TRY(upload_as(d.q_diag64, h.q_diag));
    HIP_TRY(upload_as(d.x_cv64, h.x_cv));
    HIP_TRY(upload_as(d.r_v64, h.r_v));
  }
  d.nchunks = h.nchunks;
  d.rounds = h.rounds;
  d.est_us = h.est_us;
  return RCHOL_AMD_OK;
}

// Slot-order layout of a sweep on the device from the host's natural-order
// rows: slot s takes row slot_row[s] (one thread per slot).  The host built
// rp from the same rows' lengths, so every copy stays inside both layouts.
template <typename VT>
__global__ void k_gather_slots(int n, const int *__restrict__ slot_row, const int *__restrict__ rp,
                               const int *__restrict__ nat_rp, const int *__restrict__ nat_ci,
                               const VT *__restrict__ nat_cv, const VT *__restrict__ nat_diag,
                               int *__restrict__ ci, VT *__restrict__ cv, VT *__restrict__ diag) {
  const int stride = gridDim.x * blockDim.x;
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += stride) {
    const int r = slot_row[s];
    const int a = nat_rp[r], k = nat_rp[r + 1] - a, o = rp[s];
    for (int u = 0; u < k; u++) {
      ci[o + u] = nat_ci[a + u];
      cv[o + u] = nat_cv[a + u];
    }
    diag[s] = nat_diag[r];
  }
}

// Entries and pivots of one sweep to the device: the natural-order rows
// (values rounded to VT on the host) to temporaries, gathered there into
// slot order.  sw.slot_row, sw.rp and sw.ci are already in place.
template <typename VT>
int gather_slots(rchol_amd_ctx *c, Sweep &sw, const HostSweep &h, DevBuf<VT> &cv, DevBuf<VT> &dg) {
  const size_t n = h.slot_row.size(), m = (size_t)h.rp[n];
  DevBuf<int> nrp, nci;
  DevBuf<VT> ncv, ndg;
  HIP_TRY(nrp.upload(h.nat_rp.data(), n + 1));
  HIP_TRY(nci.upload(h.nat_ci, m));
  if constexpr (sizeof(VT) == sizeof(double)) {
    HIP_TRY(ncv.upload(h.nat_cv, m));
    HIP_TRY(ndg.upload(h.nat_diag->data(), n));
  } else {
    std::vector<VT, NoInit<VT>> f(m), d(n);
    par_for(m, [&](size_t e0, size_t e1) {
      for (size_t e = e0; e < e1; e++) f[e] = (VT)h.nat_cv[e];
    });
    for (size_t r = 0; r < n; r++) d[r] = (VT)(*h.nat_diag)[r];
    HIP_TRY(ncv.upload(f.data(), m));
    HIP_TRY(ndg.upload(d.data(), n));
  }
  HIP_TRY(cv.alloc(m));
  HIP_TRY(dg.alloc(n));
  const int grid = (int)std::min<size_t>((n + kBlock - 1) / kBlock, (size_t)c->ncu * 16);
  hipLaunchKernelGGL(k_gather_slots<VT>, dim3(std::max(grid, 1)), dim3(kBlock), 0, c->stream, (int)n,
                     sw.slot_row.p, sw.rp.p, nrp.p, nci.p, ncv.p, ndg.p, sw.ci.p, cv.p, dg.p);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(c->stream));
  return RCHOL_AMD_OK;
}

// Dynamic LDS and grid of the chunked sweeps of one precision: as many
// workgroups as fit the CUs (any grid completes: the chunk queue needs no
// co-residency).
template <typename VT, typename XT>
int chunk_launch_setup(rchol_amd_ctx *c, size_t lds) {
  const void *fns[3] = {(const void *)k_sptrsv_chunk<VT, XT, double, 1>,
                        (const void *)k_sptrsv_chunk<VT, XT, XT, 2>,
                        (const void *)k_sptrsv_chunk<VT, XT, XT, 0>};
  int per_cu = 1 << 20;
  for (const void *f : fns) {
    HIP_TRY(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int b = 0;
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, f, kChunkThreads, lds));
    per_cu = std::min(per_cu, b);
  }
  if (per_cu < 1)
    return fail(RCHOL_AMD_EHIP, "the chunked sweep does not fit a CU with %zu bytes of LDS", lds);
  c->chunk_grid = per_cu * c->ncu;
  c->chunk_lds = lds;
  return RCHOL_AMD_OK;
}

// Plan (and, when chosen, upload) the chunked sweeps of a factor whose
// analysis is in hf / hb.
int setup_chunks(rchol_amd_ctx *c, size_t n, HostSweep &hf, HostSweep &hb, int nlf,
                 int nlb, int precision) {
  c->fwd.ch.reset();
  c->bwd.ch.reset();
  c->use_chunks = false;
  c->chunk_auto = false;
  const int nl = std::max(nlf, nlb);
  const bool deep = nl >= kChunkMinLevels && (double)n / nl < kChunkMaxMeanWidth;
  if (!(c->chunk_mode == 1 || (c->chunk_mode == -1 && deep))) return RCHOL_AMD_OK;
  const size_t xt = precision == RCHOL_AMD_FP32 ? sizeof(float) : sizeof(double);
  const size_t budget = c->lds_max > 2048 ? (size_t)c->lds_max - 1024 : 0;
  const int S = (int)std::min<size_t>(kChunkMaxRows, budget / xt - 1);
  if (S < kChunkMinRows)
    return c->chunk_mode == 1 ? fail(RCHOL_AMD_EHIP, "%d bytes of LDS per workgroup is too little "
                                     "for the chunked sweeps", c->lds_max)
                              : RCHOL_AMD_OK;
  const std::vector<int> bnd = chunk_bounds(hf, (int)n, S);
  HostChunks pf, pb;
  plan_chunks(hf, (int)n, bnd, false, pf);
  plan_chunks(hb, (int)n, bnd, true, pb);
  c->chunk_auto = pf.est_us + pb.est_us < (nlf + nlb) * kLevelUs;
  if (c->chunk_mode == -1 && !c->chunk_auto) return RCHOL_AMD_OK;
  int rc = upload_chunks(c->fwd.ch, pf, precision);
  if (!rc) rc = upload_chunks(c->bwd.ch, pb, precision);
  if (!rc)
    rc = precision == RCHOL_AMD_FP32 ? chunk_launch_setup<float, float>(c, (S + 1) * xt)
                                     : chunk_launch_setup<double, double>(c, (S + 1) * xt);
  if (rc) {
    c->fwd.ch.reset();
    c->bwd.ch.reset();
    return rc;
  }
  c->use_chunks = true;
  return RCHOL_AMD_OK;
}

}  // namespace

extern "C" {

int rchol_amd_upload_G(rchol_amd_ctx *c, size_t n, const size_t *colPtr, const size_t *rowIdx,
                       const double *val, int precision) {
  if (!c || !colPtr || !rowIdx || !val) return fail(RCHOL_AMD_EINVAL, "null argument");
  if (precision != RCHOL_AMD_FP64 && precision != RCHOL_AMD_FP32)
    return fail(RCHOL_AMD_EINVAL, "unknown precision %d", precision);
  if (n == 0 || n > kIdxMax) return fail(RCHOL_AMD_ERANGE, "n=%zu outside [1, 2^31)", n);
  const size_t nnz = colPtr[n];
  if (colPtr[0] != 0) return fail(RCHOL_AMD_EINVAL, "colPtr[0] must be 0");
  if (nnz > kIdxMax) return fail(RCHOL_AMD_ERANGE, "nnz(G)=%zu exceeds the int32 index range", nnz);
  HostSweep hf, hb;
  int nlf = 0, nlb = 0;
  const double t0 = host_ms();
  int rc0 = analyze_factor(n, colPtr, rowIdx, val, hf, hb, nlf, nlb);
  if (rc0) return rc0;
  const double t1 = host_ms();
  const size_t noff = nnz - n;
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->drop_graph();
  c->hasG = false;
  auto up = [&](Sweep &sw, HostSweep &h) -> int {
    sw.nslots = (int)n;
    sw.ntasks = (int)h.task_ptr.size() - 1;
    sw.nlevels = h.nlevels;
    sw.nnz_off = noff;
    HIP_TRY(sw.task_ptr.upload(h.task_ptr.data(), h.task_ptr.size()));
    sw.level_ptr_h = h.level_ptr;
    sw.rp_h.assign(h.rp.begin(), h.rp.end());
    sw.proxy_h.assign(h.proxy.begin(), h.proxy.end());
    HIP_TRY(sw.level_ptr.upload(h.level_ptr.data(), h.level_ptr.size()));
    HIP_TRY(sw.slot_row.upload(h.slot_row.data(), n));
    HIP_TRY(sw.rp.upload(h.rp.data(), n + 1));
    sw.cv64.reset();
    sw.diag64.reset();
    sw.cv32.reset();
    sw.diag32.reset();
    if (h.rp[n] != (int)noff || h.nat_rp[n] != (int)noff)
      return fail(RCHOL_AMD_EINVAL, "factor analysis lost entries");
    HIP_TRY(sw.ci.alloc(noff));
    return precision == RCHOL_AMD_FP32 ? gather_slots<float>(c, sw, h, sw.cv32, sw.diag32)
                                       : gather_slots<double>(c, sw, h, sw.cv64, sw.diag64);
    return RCHOL_AMD_OK;
  };
  int rc = up(c->fwd, hf);
  if (rc) return rc;
  rc = up(c->bwd, hb);
  if (rc) return rc;
  c->nG = n;
  c->nnzG = nnz;
  c->gprec = precision;
  c->g_gen++;
  const double t2 = host_ms();
  rc = c->size_trsv_grid();
  if (rc) return rc;
  rc = setup_chunks(c, n, hf, hb, nlf, nlb, precision);
  if (rc) return rc;
  c->hasG = true;
  c->stats.chunked = c->use_chunks ? 1 : 0;
  c->stats.chunks[0] = c->fwd.ch.nchunks;
  c->stats.chunks[1] = c->bwd.ch.nchunks;
  c->stats.chunk_rounds[0] = c->fwd.ch.rounds;
  c->stats.chunk_rounds[1] = c->bwd.ch.rounds;
  c->stats.setup_ms[1] = t1 - t0;
  c->stats.setup_ms[2] = host_ms() - t2;
  c->stats.setup_ms[3] = t2 - t1;
  c->stats.nnz_g = nnz;
  c->stats.fwd_levels = nlf;
  c->stats.bwd_levels = nlb;
  c->stats.g_precision = precision;
  return RCHOL_AMD_OK;
}

int rchol_amd_set_option(rchol_amd_ctx *c, int key, double value) {
  if (!c) return fail(RCHOL_AMD_EINVAL, "null context");
  int v = (int)value;
  switch (key) {
    case RCHOL_AMD_OPT_GRAPH_ITERS:
      if (v < 0 || v > 1024) return fail(RCHOL_AMD_EINVAL, "graph iterations %d out of [0,1024]", v);
      c->graph_iters = v;
      break;
    case RCHOL_AMD_OPT_RESID_REPLACE:
      if (v < 0) return fail(RCHOL_AMD_EINVAL, "negative replacement period");
      c->resid_period = v;
      break;
    case RCHOL_AMD_OPT_SPMV_LANES:
      if (v != 0 && v != 4 && v != 8 && v != 16 && v != 32)
        return fail(RCHOL_AMD_EINVAL, "SpMV lanes must be 0 (stream), 4, 8, 16 or 32");
      c->spmv_lanes = v;
      break;
    case RCHOL_AMD_OPT_TRSV_LANES:
      if (v < 1 || v > 64 || (v & (v - 1)))
        return fail(RCHOL_AMD_EINVAL, "sweep lanes must be a power of two in [1, 64]");
      c->trsv_lanes = v;
      break;
    case RCHOL_AMD_OPT_PROFILE:
      c->profile = v ? 1 : 0;
      break;
    case RCHOL_AMD_OPT_COHERENT_GATHER:
      c->coherent_gather = v ? 1 : 0;
      break;
    case RCHOL_AMD_OPT_STRICT_SYNC:
      c->strict_sync = v ? 1 : 0;
      break;
    case RCHOL_AMD_OPT_WAIT_TICKS: {
      if (!(value >= 0 && value <= 1e12))
        return fail(RCHOL_AMD_EINVAL, "wait bound %g not in [0, 1e12] ticks", value);
      const long long t = value == 0 ? kWaitTicks : (long long)value;
      HIP_TRY(hipSetDevice(c->device));
      HIP_TRY(hipStreamSynchronize(c->stream));
      HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(g_wait_ticks), &t, sizeof t));
      break;
    }
    case RCHOL_AMD_OPT_CHUNKED:
      if (v < -1 || v > 1) return fail(RCHOL_AMD_EINVAL, "chunked mode %d not in {-1, 0, 1}", v);
      if (c->hasG && v == 1 && c->fwd.ch.nchunks == 0)
        return fail(RCHOL_AMD_ESTATE,
                    "this factor has no chunked plan: set the option to 1 before upload_G");
      c->chunk_mode = v;
      if (c->hasG) {
        c->use_chunks = c->fwd.ch.nchunks > 0 && (v == 1 || (v == -1 && c->chunk_auto));
        c->stats.chunked = c->use_chunks ? 1 : 0;
        c->drop_graph();
      }
      break;
    case RCHOL_AMD_OPT_TASK_WAVES:
      if (v < 0) return fail(RCHOL_AMD_EINVAL, "negative task wave count");
      c->task_waves = v;
      break;
    case RCHOL_AMD_OPT_SERIAL_WIDTH:
      if (v < 0 || v > 1 << 20) return fail(RCHOL_AMD_EINVAL, "serial width %d out of range", v);
      c->serial_width = v;
      break;
    default:
      return fail(RCHOL_AMD_EINVAL, "unknown option key %d", key);
  }
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  c->drop_graph();
  if ((key == RCHOL_AMD_OPT_TRSV_LANES || key == RCHOL_AMD_OPT_SERIAL_WIDTH ||
       key == RCHOL_AMD_OPT_TASK_WAVES) && c->hasG)
    return c->size_trsv_grid();
  return RCHOL_AMD_OK;
}

int rchol_amd_get_stats(rchol_amd_ctx *c, rchol_amd_stats *out) {
  if (!c || !out) return fail(RCHOL_AMD_EINVAL, "null argument");
  *out = c->stats;
  return RCHOL_AMD_OK;
}

namespace {
// Gather (in -> staging, permuted numbering) or scatter (staging -> out).
int permute_rows(rchol_amd_ctx *c, bool gather, const double *in, double *out, int K) {
  const size_t total = c->nG * (size_t)K;
  const int g = (int)std::max<size_t>(1, std::min<size_t>((total + kBlock - 1) / kBlock, (size_t)c->ncu * 8));
  if (gather)
    hipLaunchKernelGGL(k_permute<1>, dim3(g), dim3(kBlock), 0, c->stream, (int)c->nG, K,
                       (const int *)c->perm.p, in, out);
  else
    hipLaunchKernelGGL(k_permute<0>, dim3(g), dim3(kBlock), 0, c->stream, (int)c->nG, K,
                       (const int *)c->perm.p, in, out);
  HIP_TRY(hipGetLastError());
  return RCHOL_AMD_OK;
}

int ensure_staging(rchol_amd_ctx *c, size_t m) {
  if (c->pin.n < m) HIP_TRY(c->pin.alloc(m));
  if (c->pout.n < m) HIP_TRY(c->pout.alloc(m));
  return RCHOL_AMD_OK;
}
}  // namespace

int rchol_amd_precond_device(rchol_amd_ctx *c, const double *d_r, double *d_z) {
  if (!c || !d_r || !d_z) return fail(RCHOL_AMD_EINVAL, "null argument");
  HIP_TRY(hipSetDevice(c->device));
  if (!c->hasPerm) return c->precond(d_r, d_z);
  int rc = ensure_staging(c, c->nG);
  if (!rc) rc = permute_rows(c, true, d_r, c->pin.p, 1);
  if (!rc) rc = c->precond(c->pin.p, c->pout.p);
  if (!rc) rc = permute_rows(c, false, c->pout.p, d_z, 1);
  if (!rc) HIP_TRY(hipStreamSynchronize(c->stream));
  return rc;
}

int rchol_amd_precond_device_noalloc(rchol_amd_ctx *c, const double *d_r, double *d_z) {
  if (!c || !d_r || !d_z) return fail(RCHOL_AMD_EINVAL, "null argument");
  if (!c->hasG) return fail(RCHOL_AMD_ESTATE, "G must be uploaded before precond");
  const size_t n = c->nG;
  if (c->hasA && c->n != n)
    return fail(RCHOL_AMD_EINVAL, "G is %zu x %zu but A is %zu x %zu", n, n, c->n, c->n);
  const bool fp32 = c->gprec == RCHOL_AMD_FP32;
  const ScratchLayout L = scratch_layout(n, fp32, c->hasPerm);
  if (L.bytes > kScratchBytes) {
    const size_t per_row = 8 + (c->hasPerm ? 16 : 0);  // y (+ z in FP32) + staging
    return fail(RCHOL_AMD_ERANGE,
                "n=%zu needs %zu bytes of scratch; the allocation-free entry point holds %zu "
                "(at most about %zu rows for this factor)",
                n, L.bytes, kScratchBytes, (kScratchBytes - kScratchCtl - 1024) / per_row);
  }
  if (!c->hasPerm && !fp32 && d_z < d_r + n && d_r < d_z + n)
    return fail(RCHOL_AMD_EINVAL, "d_r and d_z overlap (the FP64 sweeps write z in place)");
  HIP_TRY(hipSetDevice(c->device));
  unsigned char *base = g_scratch_dev[c->device].load();
  if (!base) return fail(RCHOL_AMD_ESTATE, "scratch not initialised on device %d", c->device);
  DevScalars *ctl = (DevScalars *)base;
  hipLaunchKernelGGL(k_clear_ctl, dim3(1), dim3(64), 0, c->stream, (unsigned int *)ctl,
                     (int)(sizeof(DevScalars) / sizeof(unsigned int)));
  const double *rhs = d_r;
  double *zout = d_z;
  if (c->hasPerm) {
    double *pin = (double *)(base + L.pin), *pout = (double *)(base + L.pout);
    int rc = permute_rows(c, true, d_r, pin, 1);
    if (rc) return rc;
    rhs = pin;
    zout = pout;
  }
  const int g = (int)std::max<size_t>(
      1, std::min<size_t>((n + kBlock - 1) / kBlock, (size_t)c->ncu * kVecBlocksPerCu));
  if (fp32) {
    float *y = (float *)(base + L.y), *z = (float *)(base + L.z);
    hipLaunchKernelGGL((k_fill_sentinel<float>), dim3(g), dim3(kBlock), 0, c->stream, (int)n, y);
    dispatch_sweeps<float, float>(c, rhs, y, z, &ctl->zero, false, nullptr, nullptr, ctl);
    hipLaunchKernelGGL(k_widen, dim3(g), dim3(kBlock), 0, c->stream, (int)n, (const float *)z,
                       zout);
  } else {
    double *y = (double *)(base + L.y);
    hipLaunchKernelGGL((k_fill_sentinel<double>), dim3(g), dim3(kBlock), 0, c->stream, (int)n, y);
    dispatch_sweeps<double, double>(c, rhs, y, zout, &ctl->zero, false, nullptr, nullptr, ctl);
  }
  if (c->hasPerm) {
    int rc = permute_rows(c, false, zout, d_z, 1);
    if (rc) return rc;
  }
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(&c->hsc->err, &ctl->err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (c->hsc->err)
    return stall(c->hsc->err);
  return RCHOL_AMD_OK;
}

int rchol_amd_set_perm(rchol_amd_ctx *c, size_t n, const size_t *perm) {
  if (!c) return fail(RCHOL_AMD_EINVAL, "null context");
  if (!perm) {
    c->hasPerm = false;
    c->perm.reset();
    return RCHOL_AMD_OK;
  }
  if (n == 0 || n > kIdxMax) return fail(RCHOL_AMD_ERANGE, "n=%zu outside [1, 2^31)", n);
  std::vector<int> p(n);
  std::vector<char> seen(n, 0);
  for (size_t i = 0; i < n; i++) {
    if (perm[i] >= n || seen[perm[i]]) return fail(RCHOL_AMD_EINVAL, "perm is not a permutation of 0..%zu", n - 1);
    seen[perm[i]] = 1;
    p[i] = (int)perm[i];
  }
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(c->perm.upload(p.data(), n));
  c->hasPerm = true;
  return RCHOL_AMD_OK;
}

int rchol_amd_edge_grad_device(rchol_amd_ctx *c, size_t m, const long long *d_u,
                               const long long *d_v, const double *d_lam, const double *d_x,
                               double *d_g) {
  if (!c || (m && (!d_u || !d_v || !d_lam || !d_x || !d_g))) return fail(RCHOL_AMD_EINVAL, "null argument");
  if (m == 0) return RCHOL_AMD_OK;
  HIP_TRY(hipSetDevice(c->device));
  const int g = (int)std::max<size_t>(1, std::min<size_t>((m + kBlock - 1) / kBlock, (size_t)c->ncu * 8));
  hipLaunchKernelGGL(k_edge_grad, dim3(g), dim3(kBlock), 0, c->stream, m, d_u, d_v, d_lam, d_x, d_g);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(c->stream));
  return RCHOL_AMD_OK;
}

int rchol_amd_precond(rchol_amd_ctx *c, const double *r, double *z) {
  if (!c || !r || !z) return fail(RCHOL_AMD_EINVAL, "null argument");
  if (!c->hasG) return fail(RCHOL_AMD_ESTATE, "G must be uploaded before precond");
  HIP_TRY(hipSetDevice(c->device));
  const size_t n = c->nG;
  DevBuf<double> dr, dz;
  HIP_TRY(dr.upload(r, n));
  HIP_TRY(dz.alloc(n));
  int rc = rchol_amd_precond_device(c, dr.p, dz.p);
  if (rc) return rc;
  HIP_TRY(hipMemcpy(z, dz.p, n * sizeof(double), hipMemcpyDeviceToHost));
  return RCHOL_AMD_OK;
}

int rchol_amd_spmv(rchol_amd_ctx *c, const double *xh, double *yh) {
  if (!c || !xh || !yh) return fail(RCHOL_AMD_EINVAL, "null argument");
  if (!c->hasA) return fail(RCHOL_AMD_ESTATE, "A must be uploaded before spmv");
  HIP_TRY(hipSetDevice(c->device));
  int rc = c->ensure_work();
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(c->tmp64a.p, xh, c->n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  launch_spmv_res(c, c->tmp64a.p, nullptr, c->q.p);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(yh, c->q.p, c->n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return RCHOL_AMD_OK;
}

int rchol_amd_pcg_device(rchol_amd_ctx *c, const double *d_b, double tol, int maxit, double *d_x,
                         int x_in_is_initial, double *relres, int *itr) {
  if (!c || !d_b || !d_x || !relres || !itr) return fail(RCHOL_AMD_EINVAL, "null argument");
  HIP_TRY(hipSetDevice(c->device));
  if (!c->hasPerm) return c->pcg(d_b, tol, maxit, d_x, x_in_is_initial, relres, itr);
  if (c->nG != c->n) return fail(RCHOL_AMD_EINVAL, "perm size does not match A");
  int rc = ensure_staging(c, c->n);
  if (!rc) rc = permute_rows(c, true, d_b, c->pin.p, 1);
  if (!rc && x_in_is_initial) rc = permute_rows(c, true, d_x, c->pout.p, 1);
  if (!rc) rc = c->pcg(c->pin.p, tol, maxit, c->pout.p, x_in_is_initial, relres, itr);
  if (!rc) rc = permute_rows(c, false, c->pout.p, d_x, 1);
  if (!rc) HIP_TRY(hipStreamSynchronize(c->stream));
  return rc;
}

int rchol_amd_pcg(rchol_amd_ctx *c, const double *bh, double tol, int maxit, double *xh,
                  int x_in_is_initial, double *relres, int *itr) {
  if (!c || !bh || !xh || !relres || !itr) return fail(RCHOL_AMD_EINVAL, "null argument");
  if (!c->hasA) return fail(RCHOL_AMD_ESTATE, "A must be uploaded before pcg");
  HIP_TRY(hipSetDevice(c->device));
  int rc = c->ensure_work();
  if (rc) return rc;
  if (c->hasPerm) {
    DevBuf<double> db, dx;
    HIP_TRY(db.upload(bh, c->n));
    if (x_in_is_initial)
      HIP_TRY(dx.upload(xh, c->n));
    else
      HIP_TRY(dx.alloc(c->n));
    rc = rchol_amd_pcg_device(c, db.p, tol, maxit, dx.p, x_in_is_initial, relres, itr);
    if (rc) return rc;
    HIP_TRY(hipMemcpy(xh, dx.p, c->n * sizeof(double), hipMemcpyDeviceToHost));
    return RCHOL_AMD_OK;
  }
  HIP_TRY(hipMemcpyAsync(c->b.p, bh, c->n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  if (x_in_is_initial)
    HIP_TRY(hipMemcpyAsync(c->x.p, xh, c->n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  rc = c->pcg(c->b.p, tol, maxit, c->x.p, x_in_is_initial, relres, itr);
  if (rc) return rc;
  HIP_TRY(hipMemcpy(xh, c->x.p, c->n * sizeof(double), hipMemcpyDeviceToHost));
  return RCHOL_AMD_OK;
}


int rchol_amd_pcg_batch(rchol_amd_ctx *c, int k, const double *B, double tol, int maxit,
                        double *X, int x_in_is_initial, double *relres, int *itr) {
  if (!c || !B || !X || !relres || !itr) return fail(RCHOL_AMD_EINVAL, "null argument");
  if (k < 1 || k > kMaxCols) return fail(RCHOL_AMD_EINVAL, "batch width %d outside [1, %d]", k, kMaxCols);
  if (!c->hasA) return fail(RCHOL_AMD_ESTATE, "A must be uploaded before pcg");
  HIP_TRY(hipSetDevice(c->device));
  const int K = batch_width(k);
  const size_t n = c->n, m = n * (size_t)K;
  int rc = c->ensure_batch(K);
  if (rc) return rc;
  // pack the caller's n x k interleaved block into n x K
  std::vector<double> hb(m, 0.0), hx;
  std::vector<size_t> P(n);
  if (c->hasPerm) {
    std::vector<int> hp(n);
    HIP_TRY(hipMemcpy(hp.data(), c->perm.p, n * sizeof(int), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; i++) P[i] = (size_t)hp[i];
  } else {
    for (size_t i = 0; i < n; i++) P[i] = i;
  }
  for (size_t i = 0; i < n; i++)
    for (int j = 0; j < k; j++) hb[i * K + j] = B[P[i] * k + j];
  HIP_TRY(hipMemcpyAsync(c->batch.b.p, hb.data(), m * sizeof(double), hipMemcpyHostToDevice, c->stream));
  if (x_in_is_initial) {
    hx.assign(m, 0.0);
    for (size_t i = 0; i < n; i++)
      for (int j = 0; j < k; j++) hx[i * K + j] = X[P[i] * k + j];
    HIP_TRY(hipMemcpyAsync(c->batch.x.p, hx.data(), m * sizeof(double), hipMemcpyHostToDevice, c->stream));
  }
  std::vector<double> rr(kMaxCols);
  std::vector<int> it(kMaxCols);
  rc = pcg_batch_dev(c, K, c->batch.b.p, k, tol, maxit, c->batch.x.p, x_in_is_initial, rr.data(), it.data());
  if (rc) return rc;
  hx.assign(m, 0.0);
  HIP_TRY(hipMemcpy(hx.data(), c->batch.x.p, m * sizeof(double), hipMemcpyDeviceToHost));
  for (size_t i = 0; i < n; i++)
    for (int j = 0; j < k; j++) X[P[i] * k + j] = hx[i * K + j];
  for (int j = 0; j < k; j++) {
    relres[j] = rr[j];
    itr[j] = it[j];
  }
  return RCHOL_AMD_OK;
}

int rchol_amd_pcg_batch_device(rchol_amd_ctx *c, int k, const double *d_B, double tol, int maxit,
                               double *d_X, int x_in_is_initial, double *relres, int *itr) {
  if (!c || !d_B || !d_X || !relres || !itr) return fail(RCHOL_AMD_EINVAL, "null argument");
  if (k != 4 && k != 8 && k != 16 && k != 32)
    return fail(RCHOL_AMD_EINVAL, "device batches take 4, 8, 16 or 32 columns (got %d)", k);
  HIP_TRY(hipSetDevice(c->device));
  if (!c->hasPerm) return pcg_batch_dev(c, k, d_B, k, tol, maxit, d_X, x_in_is_initial, relres, itr);
  const size_t m = c->n * (size_t)k;
  int rc = ensure_staging(c, m);
  if (!rc) rc = permute_rows(c, true, d_B, c->pin.p, k);
  if (!rc && x_in_is_initial) rc = permute_rows(c, true, d_X, c->pout.p, k);
  if (!rc) rc = pcg_batch_dev(c, k, c->pin.p, k, tol, maxit, c->pout.p, x_in_is_initial, relres, itr);
  if (!rc) rc = permute_rows(c, false, c->pout.p, d_X, k);
  if (!rc) HIP_TRY(hipStreamSynchronize(c->stream));
  return rc;
}

int rchol_amd_precond_batch(rchol_amd_ctx *c, int k, const double *R, double *Z) {
  if (!c || !R || !Z) return fail(RCHOL_AMD_EINVAL, "null argument");
  if (k < 1 || k > kMaxCols) return fail(RCHOL_AMD_EINVAL, "batch width %d outside [1, %d]", k, kMaxCols);
  if (!c->hasG) return fail(RCHOL_AMD_ESTATE, "G must be uploaded before precond");
  HIP_TRY(hipSetDevice(c->device));
  if (!c->hasA) c->n = c->nG;
  const int K = batch_width(k);
  const size_t n = c->nG, m = n * (size_t)K;
  int rc = c->ensure_batch(K);
  if (rc) return rc;
  std::vector<double> h(m, 0.0);
  std::vector<size_t> P(n);
  if (c->hasPerm) {
    std::vector<int> hp(n);
    HIP_TRY(hipMemcpy(hp.data(), c->perm.p, n * sizeof(int), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; i++) P[i] = (size_t)hp[i];
  } else {
    for (size_t i = 0; i < n; i++) P[i] = i;
  }
  for (size_t i = 0; i < n; i++)
    for (int j = 0; j < k; j++) h[i * K + j] = R[P[i] * k + j];
  HIP_TRY(hipMemcpy(c->batch.b.p, h.data(), m * sizeof(double), hipMemcpyHostToDevice));
  rc = precond_batch_dev(c, K, c->batch.b.p, c->batch.res.p);
  if (rc) return rc;
  HIP_TRY(hipMemcpy(h.data(), c->batch.res.p, m * sizeof(double), hipMemcpyDeviceToHost));
  for (size_t i = 0; i < n; i++)
    for (int j = 0; j < k; j++) Z[P[i] * k + j] = h[i * K + j];
  return RCHOL_AMD_OK;
}


int rchol_amd_set_segments(rchol_amd_ctx *c, size_t nseg, const size_t *off) {
  if (!c || !off) return fail(RCHOL_AMD_EINVAL, "null argument");
  if (nseg < 1 || nseg > (size_t)kMaxSeg)
    return fail(RCHOL_AMD_EINVAL, "segment count %zu outside [1, %d]", nseg, kMaxSeg);
  if (!c->hasA) return fail(RCHOL_AMD_ESTATE, "A must be uploaded before the segments");
  if (off[0] != 0 || off[nseg] != c->n) return fail(RCHOL_AMD_EINVAL, "segments must cover 0..n");
  for (size_t s = 0; s < nseg; s++)
    if (off[s + 1] <= off[s]) return fail(RCHOL_AMD_EINVAL, "segment %zu is empty", s);
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->segs.drop_graph();
  // segment-major block layout, ~ncu*8 blocks in proportion to rows
  const size_t target = (size_t)c->ncu * 8;
  std::vector<int> seg, r0, r1, sb(1, 0);
  for (size_t s = 0; s < nseg; s++) {
    const size_t rows = off[s + 1] - off[s];
    size_t nb = std::max<size_t>(1, (size_t)std::llround((double)target * rows / c->n));
    nb = std::min(nb, (rows + 255) / 256);
    nb = std::max<size_t>(nb, 1);
    for (size_t b = 0; b < nb; b++) {
      seg.push_back((int)s);
      r0.push_back((int)(off[s] + rows * b / nb));
      r1.push_back((int)(off[s] + rows * (b + 1) / nb));
    }
    sb.push_back((int)seg.size());
  }
  HIP_TRY(c->segs.seg.upload(seg.data(), seg.size()));
  HIP_TRY(c->segs.r0.upload(r0.data(), r0.size()));
  HIP_TRY(c->segs.r1.upload(r1.data(), r1.size()));
  HIP_TRY(c->segs.sb.upload(sb.data(), sb.size()));
  if (!c->segs.ss) {
    HIP_TRY(hipMalloc((void **)&c->segs.ss, sizeof(SegScalars)));
    HIP_TRY(hipHostMalloc((void **)&c->segs.hss, sizeof(SegScalars), hipHostMallocDefault));
  }
  HIP_TRY(hipMemset(c->segs.ss, 0, sizeof(SegScalars)));
  memset(c->segs.hss, 0, sizeof(SegScalars));
  c->segs.S = (int)nseg;
  c->segs.nblocks = (int)seg.size();
  c->segs.off.assign(off, off + nseg + 1);
  // partial sums: 2 per block
  if (c->partials.n < (size_t)c->segs.nblocks * 2 + 64)
    HIP_TRY(c->partials.alloc((size_t)c->segs.nblocks * 2 + 64));
  return RCHOL_AMD_OK;
}

int rchol_amd_pcg_segments(rchol_amd_ctx *c, const double *b, double tol, int maxit, double *x,
                           int x_in_is_initial, double *relres, int *itr) {
  if (!c || !b || !x || !relres || !itr) return fail(RCHOL_AMD_EINVAL, "null argument");
  HIP_TRY(hipSetDevice(c->device));
  int rc = c->ensure_work();
  if (rc) return rc;
  if (c->partials.n < (size_t)c->segs.nblocks * 2 + 64)
    HIP_TRY(c->partials.alloc((size_t)c->segs.nblocks * 2 + 64));
  HIP_TRY(hipMemcpyAsync(c->b.p, b, c->n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  if (x_in_is_initial)
    HIP_TRY(hipMemcpyAsync(c->x.p, x, c->n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  rc = c->pcg_segments(c->b.p, tol, maxit, c->x.p, x_in_is_initial, relres, itr);
  if (rc) return rc;
  HIP_TRY(hipMemcpy(x, c->x.p, c->n * sizeof(double), hipMemcpyDeviceToHost));
  return RCHOL_AMD_OK;
}


int rchol_amd_autotune(rchol_amd_ctx *c, int reps, int *trsv_lanes, int *spmv_lanes,
                       double *trsv_ms, double *spmv_ms) {
  if (!c) return fail(RCHOL_AMD_EINVAL, "null context");
  if (!c->hasA || !c->hasG) return fail(RCHOL_AMD_ESTATE, "A and G must be uploaded before tuning");
  if (reps < 1) reps = 3;
  HIP_TRY(hipSetDevice(c->device));
  int rc = c->ensure_work();
  if (rc) return rc;
  const size_t n = c->n;
  std::vector<double> h(n);
  for (size_t i = 0; i < n; i++) h[i] = 1.0 + (double)((i * 2654435761u) % 1000) * 1e-3;
  HIP_TRY(hipMemcpy(c->b.p, h.data(), n * sizeof(double), hipMemcpyHostToDevice));
  const int cand_t[4] = {1, 2, 4, 8};  // minimum lanes per sweep row
  const int cand_s[4] = {0, 4, 8, 16};  // lanes per SpMV row (0 = stream form)
  auto time_ms = [&](auto fn) -> double {
    std::vector<float> t;
    for (int r = 0; r <= reps; r++) {  // first run warms up
      (void)hipEventRecord(c->ev0, c->stream);
      fn();
      (void)hipEventRecord(c->ev1, c->stream);
      (void)hipStreamSynchronize(c->stream);
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
      if (r) t.push_back(ms);
    }
    std::sort(t.begin(), t.end());
    return t[t.size() / 2];
  };
  int best_t = c->trsv_lanes, best_s = c->spmv_lanes;
  double bt = 1e30, bs = 1e30;
  for (int L : cand_t) {
    c->trsv_lanes = L;
    rc = c->size_trsv_grid();
    if (rc) return rc;
    double ms = time_ms([&]() { (void)c->precond(c->b.p, c->tmp64a.p); });
    if (ms < bt) { bt = ms; best_t = L; }
  }
  for (int L : cand_s) {
    c->spmv_lanes = L;
    double ms = time_ms([&]() { launch_spmv_res(c, c->b.p, nullptr, c->q.p); });
    if (ms < bs) { bs = ms; best_s = L; }
  }
  c->trsv_lanes = best_t;
  c->spmv_lanes = best_s;
  rc = c->size_trsv_grid();
  if (rc) return rc;
  c->drop_graph();
  if (trsv_lanes) *trsv_lanes = best_t;
  if (spmv_lanes) *spmv_lanes = best_s;
  if (trsv_ms) *trsv_ms = bt;
  if (spmv_ms) *spmv_ms = bs;
  return RCHOL_AMD_OK;
}


int rchol_amd_copy_probe(rchol_amd_ctx *c, size_t bytes, int reps, double *gbs, double *ms) {
  if (!c || !gbs) return fail(RCHOL_AMD_EINVAL, "null argument");
  if (bytes < 4096) return fail(RCHOL_AMD_EINVAL, "copy probe needs at least 4 KiB");
  if (reps < 1) reps = 5;
  HIP_TRY(hipSetDevice(c->device));
  const size_t n16 = bytes / 16;
  DevBuf<double2> a, b;
  HIP_TRY(a.alloc(n16));
  HIP_TRY(b.alloc(n16));
  HIP_TRY(hipMemsetAsync(a.p, 0, n16 * 16, c->stream));
  const int g = c->ncu * 8;
  std::vector<float> t;
  for (int r = 0; r <= reps; r++) {
    HIP_TRY(hipEventRecord(c->ev0, c->stream));
    hipLaunchKernelGGL(k_copy, dim3(g), dim3(kBlock), 0, c->stream, n16, (const double2 *)a.p, b.p);
    HIP_TRY(hipEventRecord(c->ev1, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    float m = 0.f;
    HIP_TRY(hipEventElapsedTime(&m, c->ev0, c->ev1));
    if (r) t.push_back(m);
  }
  std::sort(t.begin(), t.end());
  const double med = t[t.size() / 2];
  *gbs = 2.0 * (double)(n16 * 16) / (med * 1e-3) / 1e9;
  if (ms) *ms = med;
  return RCHOL_AMD_OK;
}


int rchol_amd_analyze_G(size_t n, const size_t *colPtr, const size_t *rowIdx, const double *val,
                        int serial_width, rchol_amd_schedule *out) {
  if (!colPtr || !rowIdx || !val || !out) return fail(RCHOL_AMD_EINVAL, "null argument");
  if (n == 0 || n > kIdxMax) return fail(RCHOL_AMD_ERANGE, "n=%zu outside [1, 2^31)", n);
  if (colPtr[0] != 0) return fail(RCHOL_AMD_EINVAL, "colPtr[0] must be 0");
  if (colPtr[n] > kIdxMax) return fail(RCHOL_AMD_ERANGE, "nnz(G) exceeds the int32 index range");
  HostSweep h[2];
  int nl[2];
  int rc = analyze_factor(n, colPtr, rowIdx, val, h[0], h[1], nl[0], nl[1]);
  if (rc) return rc;
  memset(out, 0, sizeof(*out));
  for (int k = 0; k < 2; k++) {
    int hd, tl;
    compute_split(h[k].level_ptr, nl[k], serial_width, hd, tl);
    int tasks = 0;
    for (int l = hd; l < tl; l++) tasks += (h[k].level_ptr[l + 1] - h[k].level_ptr[l] + kTaskRows - 1) / kTaskRows;
    out->nlevels[k] = nl[k];
    out->head_end[k] = hd;
    out->tail_begin[k] = tl;
    out->mid_tasks[k] = tasks;
    out->level_ptr[k] = (int *)malloc((nl[k] + 1) * sizeof(int));
    out->slot_row[k] = (int *)malloc(n * sizeof(int));
    out->proxy[k] = (unsigned short *)malloc(n * sizeof(unsigned short));
    if (!out->level_ptr[k] || !out->slot_row[k] || !out->proxy[k])
      return fail(RCHOL_AMD_EINVAL, "out of memory");
    memcpy(out->level_ptr[k], h[k].level_ptr.data(), (nl[k] + 1) * sizeof(int));
    memcpy(out->slot_row[k], h[k].slot_row.data(), n * sizeof(int));
    memcpy(out->proxy[k], h[k].proxy.data(), n * sizeof(unsigned short));
  }
  return RCHOL_AMD_OK;
}

}  // extern "C"

namespace {

// The chunked plan of one sweep run on the host in the kernel's order of
// operations (double arithmetic), checking its invariants: a chunk reads
// only rows of chunks dequeued before it, every thread's run of external
// entries starts on a row boundary, a round reads only rows that no later
// round writes.
int simulate_chunks(const HostChunks &p, int n, const double *rhs, double *out) {
  const int C = p.nchunks, I = C + p.nparts;
  std::vector<char> done(n, 0), listed(n, 0), item_done(I, 0);
  std::vector<double> yl, extsum(n, 0.0);
  std::vector<int> last;
  // rows of the chunks an item may read: those of the chunks it waits for,
  // through its parts for a chunk with parts
  auto mark = [&](int it, char v) {
    for (int d = p.dep_ptr[it]; d < p.dep_ptr[it + 1]; d++) {
      const int cd = p.dep[d];
      if (cd >= C) continue;
      for (int l = 0; l < p.m[cd]; l++) listed[p.q_row[p.q0[cd] + l]] = v;
    }
  };
  // thread runs of external entries, summed per row in entry order
  auto runs = [&](const int *sp, int it, auto &&sink) -> int {
    for (int t = 0; t < kChunkThreads; t++) {
      if (sp[t] > sp[t + 1]) return fail(RCHOL_AMD_EINVAL, "item %d: split not monotone", it);
      if (t > 0 && sp[t] > sp[0] && sp[t] < sp[kChunkThreads] && p.x_l[sp[t]] == p.x_l[sp[t] - 1])
        return fail(RCHOL_AMD_EINVAL, "item %d: thread %d starts inside a row", it, t);
      int cur = -1;
      double acc = 0.0;
      for (int e = sp[t]; e < sp[t + 1]; e++) {
        const int j = p.x_ci[e];
        if (!done[j]) return fail(RCHOL_AMD_EINVAL, "item %d reads row %d before it is published", it, j);
        if (!listed[j])
          return fail(RCHOL_AMD_EINVAL, "item %d reads row %d of a chunk it does not wait for", it, j);
        if (p.x_l[e] != cur) {
          if (cur >= 0) sink(cur, acc);
          cur = p.x_l[e];
          acc = 0.0;
        }
        acc += p.x_cv[e] * out[j];
      }
      if (cur >= 0) sink(cur, acc);
    }
    return RCHOL_AMD_OK;
  };
  for (int k = 0; k < I; k++) {
    const int it = p.order[k];
    for (int d = p.dep_ptr[it]; d < p.dep_ptr[it + 1]; d++)
      if (!item_done[p.dep[d]])
        return fail(RCHOL_AMD_EINVAL, "item %d dequeued before item %d it waits for", it, p.dep[d]);
    if (it >= C) {  // a part: external sums of rows [l0, l1) of its chunk
      const int q = it - C, c = p.p_chunk[q], q0 = p.q0[c];
      if (p.p_lg[q] < 0 || p.p_lg[q] > 6) return fail(RCHOL_AMD_EINVAL, "part %d: lane width", q);
      mark(it, 1);
      for (int l = p.p_l0[q]; l < p.p_l1[q]; l++) {
        double acc = 0.0;
        for (int e = p.x_rp[q0 + l]; e < p.x_rp[q0 + l + 1]; e++) {
          const int j = p.x_ci[e];
          if (p.x_l[e] != l) return fail(RCHOL_AMD_EINVAL, "part %d: entry %d of another row", q, e);
          if (!done[j] || !listed[j])
            return fail(RCHOL_AMD_EINVAL, "part %d reads row %d before it is published", q, j);
          acc += p.x_cv[e] * out[j];
        }
        extsum[q0 + l] = acc;
      }
      mark(it, 0);
      item_done[it] = 1;
      continue;
    }
    const int c = it;
    const int q0 = p.q0[c], m = p.m[c];
    const bool parted = p.c_part0[c + 1] > p.c_part0[c];
    yl.assign(m, 0.0);
    for (int l = 0; l < m; l++) yl[l] = rhs[p.q_row[q0 + l]] - (parted ? extsum[q0 + l] : 0.0);
    if (parted) {  // the parts cover every row, in order
      int next = 0;
      for (int q = p.c_part0[c]; q < p.c_part0[c + 1]; q++) {
        if (p.p_l0[q] != next) return fail(RCHOL_AMD_EINVAL, "chunk %d: parts leave a gap", c);
        next = p.p_l1[q];
      }
      if (next != m) return fail(RCHOL_AMD_EINVAL, "chunk %d: parts end before its rows", c);
    }
    mark(it, 1);
    int rc = runs(&p.split[(size_t)c * (kChunkThreads + 1)], it,
                  [&](int l, double a) { yl[l] -= a; });
    mark(it, 0);
    if (rc) return rc;
    for (int l = 0; l < m; l++) yl[l] = yl[l] / p.q_diag[q0 + l];
    last.assign(m, -1);
    for (int r = p.r0[c]; r < p.r1[c]; r++)
      for (int ln = 0; ln < 64; ln++) {
        const unsigned w = p.r_w[(size_t)r * 64 + ln];
        if (w == kChunkPad) continue;
        if ((int)(w >> 16) >= m || (int)(w & 0xFFFFu) >= m)
          return fail(RCHOL_AMD_EINVAL, "chunk %d: round %d names a row outside the chunk", c, r);
        last[w >> 16] = r;
      }
    for (int r = p.r0[c]; r < p.r1[c]; r++)
      for (int ln = 0; ln < 64; ln++) {
        const unsigned w = p.r_w[(size_t)r * 64 + ln];
        if (w == kChunkPad) continue;
        const unsigned src = w & 0xFFFFu, tgt = w >> 16;
        if (last[src] >= r)
          return fail(RCHOL_AMD_EINVAL, "chunk %d: round %d reads row %u, written in round %d", c,
                      r, src, last[src]);
        yl[tgt] -= p.r_v[(size_t)r * 64 + ln] * yl[src];
      }
    for (int l = 0; l < m; l++) {
      const int row = p.q_row[q0 + l];
      if (done[row]) return fail(RCHOL_AMD_EINVAL, "row %d belongs to two chunks", row);
      out[row] = yl[l];
      done[row] = 1;
    }
    item_done[it] = 1;
  }
  for (int i = 0; i < n; i++)
    if (!done[i]) return fail(RCHOL_AMD_EINVAL, "row %d is in no chunk", i);
  return RCHOL_AMD_OK;
}

}  // namespace

extern "C" {

int rchol_amd_chunk_apply_host(size_t n, const size_t *colPtr, const size_t *rowIdx,
                               const double *val, int max_rows, const double *r, double *z,
                               int *counts, long long *rounds, double *est_us) {
  if (!colPtr || !rowIdx || !val || !r || !z) return fail(RCHOL_AMD_EINVAL, "null argument");
  if (n == 0 || n > kIdxMax) return fail(RCHOL_AMD_ERANGE, "n=%zu outside [1, 2^31)", n);
  if (colPtr[0] != 0) return fail(RCHOL_AMD_EINVAL, "colPtr[0] must be 0");
  if (colPtr[n] > kIdxMax) return fail(RCHOL_AMD_ERANGE, "nnz(G) exceeds the int32 index range");
  if (max_rows < 1 || max_rows > 65535)
    return fail(RCHOL_AMD_EINVAL, "chunk rows %d not in [1, 65535]", max_rows);
  HostSweep hf, hb;
  int nlf = 0, nlb = 0;
  int rc = analyze_factor(n, colPtr, rowIdx, val, hf, hb, nlf, nlb);
  if (rc) return rc;
  const std::vector<int> bnd = chunk_bounds(hf, (int)n, max_rows);
  HostChunks pf, pb;
  plan_chunks(hf, (int)n, bnd, false, pf);
  plan_chunks(hb, (int)n, bnd, true, pb);
  std::vector<double> y(n);
  rc = simulate_chunks(pf, (int)n, r, y.data());
  if (!rc) rc = simulate_chunks(pb, (int)n, y.data(), z);
  if (rc) return rc;
  if (counts) {
    counts[0] = pf.nchunks;
    counts[1] = pf.nparts;
    counts[2] = pb.nparts;
  }
  if (rounds) *rounds = pf.rounds + pb.rounds;
  if (est_us) {
    est_us[0] = pf.est_us;
    est_us[1] = pb.est_us;
  }
  return RCHOL_AMD_OK;
}

void rchol_amd_schedule_free(rchol_amd_schedule *s) {
  if (!s) return;
  for (int k = 0; k < 2; k++) {
    free(s->level_ptr[k]);
    free(s->slot_row[k]);
    free(s->proxy[k]);
    s->level_ptr[k] = nullptr;
    s->slot_row[k] = nullptr;
    s->proxy[k] = nullptr;
  }
}

int rchol_amd_pcg_csr(size_t n, const size_t *aRowPtr, const size_t *aColIdx, const double *aVal,
                      const double *b, double tol, int maxit, const size_t *gColPtr,
                      const size_t *gRowIdx, const double *gVal, double *x, int x_in_is_initial,
                      double *relres, int *itr) {
  rchol_amd_ctx *c = nullptr;
  int rc = rchol_amd_create(0, &c);
  if (rc) return rc;
  rc = rchol_amd_upload_A(c, n, aRowPtr, aColIdx, aVal);
  if (!rc) rc = rchol_amd_upload_G(c, n, gColPtr, gRowIdx, gVal, RCHOL_AMD_FP64);
  if (!rc) rc = rchol_amd_pcg(c, b, tol, maxit, x, x_in_is_initial, relres, itr);
  std::string keep = g_err;
  rchol_amd_destroy(c);
  g_err = keep;
  return rc;
}

}  // extern "C"
