// Plans, the time-step driver and the remaining C ABI entry points.
#include "wfs_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <vector>

struct wfs_plan {
  int n[3];
  int batch;              // independent grids stacked along x (same shape and parameters)
  double dx[3];
  double dt, mass;
  int precision, method, order;
  size_t elem;            // bytes per state element
  void* tw[3];            // forward twiddles per axis, state precision; ntab dithered copies
  int ntab;               // copies of each per-axis table (complex64: kDither, complex128: 1)
  int step_index;         // table copy of the next step, in [0, ntab); advanced by every call's steps
  void* kfull[3];         // exp(-i k_d^2 dt / 2m) / N_d, state precision, ntab copies
  void* khalf[3];         // exp(-i k_d^2 dt / 4m) / N_d, state precision, ntab copies
  void* scratch[2];       // power-series ping-pong buffers (lazy)
  double* obs_work;       // observable partial records
  long obs_work_len;
  double* plane_work;     // plane-current partial records (grown on demand)
  long plane_work_len;
  double* plane_out;      // 7 doubles
  void* pv;               // potential phase table cexp_i(-dt V), state precision, C order
  void* pv_lines;         // the same table in the slab scheme's lines order (LinesOrder)
  double* obs_lines;      // slab scheme: box moment records of an observing launch (lazy)
  long obs_lines_len;
  bool pv_nat_ok;         // pv holds the current potential
  bool pv_lines_ok;       // pv_lines holds the current potential
  int* pv_planes;         // slab scheme: per x plane the first plane of its run of identical
                          // planes of pv_lines, then as many ints of workspace (lazy)
  bool pv_planes_ok;      // pv_planes describes the current pv_lines
  bool pv_in_graph;       // pv_lines was written inside a captured graph: its content changes at
                          // replays the plan does not see, so the plane map is not used again
  bool pv_set;
  void* tw_rows[2];       // slab scheme: the y and z twiddles of each thread's lines, [ntab][u][k1]
  void* sep[3];           // separable potential: per-axis exp(-i dt f_d), [ntab][batch][n_d]
  bool sep_set;           // the step uses sep[] instead of pv
  int variant[WFS_K_COUNT];   // tile variant per kernel class (wfs_plan_autotune)
  // per-kernel timing (wfs_plan_set_profiling)
  bool profiling;
  std::vector<std::pair<hipEvent_t, hipEvent_t>>* pending;   // [WFS_K_COUNT]
  std::vector<hipEvent_t>* event_pool;
  long launches[WFS_K_COUNT];
  double total_ms[WFS_K_COUNT];
};

namespace wfs {

static thread_local char g_last_error[512] = "";

LaunchTiming*& launch_timing() {
  static thread_local LaunchTiming* current = nullptr;
  return current;
}

void set_last_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
}

namespace {

#define WFS_TRY(expr)              \
  do {                             \
    int _rc = (expr);              \
    if (_rc != 0) return _rc;      \
  } while (0)

// numpy.fft.fftfreq index -> signed frequency index
long freq_index(long j, long n) { return (j <= (n - 1) / 2) ? j : j - n; }

int upload(void** dst, const void* host, size_t bytes) {
  hipError_t e = hipMalloc(dst, bytes);
  if (e != hipSuccess) { set_last_error("hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e)); return WFS_ERR_ALLOC; }
  e = hipMemcpy(*dst, host, bytes, hipMemcpyHostToDevice);
  if (e != hipSuccess) { set_last_error("hipMemcpy failed: %s", hipGetErrorString(e)); return int(e); }
  return 0;
}

// complex table in the state precision from FP64 (re, im) pairs
int upload_table(void** dst, const std::vector<double>& re, const std::vector<double>& im, int precision) {
  const size_t n = re.size();
  if (precision == WFS_C128) {
    std::vector<double2> h(n);
    for (size_t j = 0; j < n; ++j) { h[j].x = re[j]; h[j].y = im[j]; }
    return upload(dst, h.data(), n * sizeof(double2));
  }
  std::vector<float2> h(n);
  for (size_t j = 0; j < n; ++j) { h[j].x = float(re[j]); h[j].y = float(im[j]); }
  return upload(dst, h.data(), n * sizeof(float2));
}

// Copies of a per-axis table for complex64 plans, cycled step by step.  A factor
// rounded to float and applied to the same mode every step errs the same way every
// step, so its error grows linearly with the step count.  Each part x of every entry
// is rounded down in some copies and up in the others, in proportion to where x lies
// between the two floats, so the mean over the copies is within 1/(2 kDither) ulp of x;
// which copies round up rotates with the entry, so no copy is biased as a whole.
constexpr int kDither = 8;

float dither_part(double x, long entry, int part, int copy) {
  float lo = float(x);   // nearest float, then the float at or below x
  if (double(lo) > x) lo = std::nextafter(lo, -INFINITY);
  if (double(lo) == x) return lo;
  const float hi = std::nextafter(lo, INFINITY);
  const double frac = (x - double(lo)) / (double(hi) - double(lo));
  const int ups = int(std::lround(frac * kDither));
  const int slot = int((copy + 3 * entry + 5 * part) % kDither);
  return slot < ups ? hi : lo;
}

int upload_dithered(void** dst, const std::vector<double>& re, const std::vector<double>& im, int precision) {
  if (precision == WFS_C128) return upload_table(dst, re, im, precision);
  const size_t n = re.size();
  std::vector<float2> h(n * kDither);
  for (int c = 0; c < kDither; ++c)
    for (size_t j = 0; j < n; ++j) {
      h[c * n + j].x = dither_part(re[j], long(j), 0, c);
      h[c * n + j].y = dither_part(im[j], long(j), 1, c);
    }
  return upload(dst, h.data(), h.size() * sizeof(float2));
}

// The twiddles W_L^{u k1} = W_N^{(M u k1) mod N} (u < U, k1 < K) of the length-L lines
// of the slab scheme's y-z launch, each copy laid out [u][k1] so that a thread's K twiddles
// are contiguous (wide scalar loads); entry for entry the values of the dithered N-table.
int upload_rows_table(void** dst, const std::vector<double>& re, const std::vector<double>& im, int M, int U,
                      int K) {
  const long n = long(re.size());
  std::vector<float2> h(size_t(kDither) * U * K);
  for (int c = 0; c < kDither; ++c)
    for (int u = 0; u < U; ++u)
      for (int k1 = 0; k1 < K; ++k1) {
        const long j = (long(M) * u * k1) % n;
        float2& e = h[(size_t(c) * U + u) * K + k1];
        e.x = dither_part(re[j], j, 0, c);
        e.y = dither_part(im[j], j, 1, c);
      }
  return upload(dst, h.data(), h.size() * sizeof(float2));
}

// cos/sin(2 pi j / n) with exact values at multiples of pi/2 (long double)
void unit_root(long j, long n, double& c, double& s) {
  j %= n;
  if ((4 * j) % n == 0) {
    const long q = (4 * j) / n;
    const double cq[4] = {1.0, 0.0, -1.0, 0.0}, sq[4] = {0.0, 1.0, 0.0, -1.0};
    c = cq[q]; s = sq[q];
    return;
  }
  const long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)j / (long double)n;
  c = double(cosl(a));
  s = double(sinl(a));
}

// exp(-i k^2 dt / (f m)) / n with the phase reduced modulo 2 pi in long double
void kinetic_table(long n, double dx, double dt, double mass, double f, std::vector<double>& re,
                   std::vector<double>& im) {
  re.resize(n); im.resize(n);
  const long double two_pi = 2.0L * 3.14159265358979323846264338327950288L;
  for (long j = 0; j < n; ++j) {
    const long double k = two_pi * (long double)freq_index(j, n) / ((long double)n * (long double)dx);
    long double th = -k * k * (long double)dt / ((long double)f * (long double)mass);
    th = fmodl(th, two_pi);
    re[j] = double(cosl(th) / (long double)n);
    im[j] = double(sinl(th) / (long double)n);
  }
}

// Lazy allocations cannot be recorded into a graph: a call that would allocate on a stream under
// capture fails with a message instead (the caller makes one uncaptured call first).
// stream s is being captured into a graph (its launches run only at replay)
bool capturing(hipStream_t s) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  return s && hipStreamIsCapturing(s, &st) == hipSuccess && st != hipStreamCaptureStatusNone;
}

int refuse_if_capturing(hipStream_t s, const char* what) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  if (s && hipStreamIsCapturing(s, &st) == hipSuccess && st != hipStreamCaptureStatusNone) {
    set_last_error("%s: its first call allocates a workspace and cannot be captured; make one uncaptured call first",
                   what);
    return WFS_ERR_INVALID;
  }
  return 0;
}

int ensure_scratch(wfs_plan* p, hipStream_t s) {
  if (p->scratch[0]) return 0;
  WFS_TRY(refuse_if_capturing(s, "power-series wfs_step"));
  const size_t bytes = size_t(p->batch) * p->n[0] * p->n[1] * p->n[2] * p->elem;
  for (int i = 0; i < 2; ++i) {
    hipError_t e = hipMalloc(&p->scratch[i], bytes);
    if (e != hipSuccess) { set_last_error("hipMalloc(%zu) for power-series scratch failed", bytes); return WFS_ERR_ALLOC; }
  }
  return 0;
}

// Timing of the launches of one kernel class when profiling is on: the launches made through
// WFS_LAUNCH inside the scope record the kernels' own start and end timestamps into the event
// pair (the durations a kernel trace reports).  On a stream under capture the pair is recorded
// around the launches instead, as graph nodes.
struct ProfileScope {
  wfs_plan* p;
  int kclass;
  hipStream_t s;
  LaunchTiming t;
  LaunchTiming* outer = nullptr;
  bool hooked = false;
  ProfileScope(wfs_plan* plan, int kc, hipStream_t st) : p(plan), kclass(kc), s(st) {
    if (!p->profiling) return;
    t.start = take();
    t.stop = take();
    if (!t.start || !t.stop) return;
    if (capturing(s)) {
      (void)hipEventRecord(t.start, s);
    } else {
      outer = launch_timing();
      launch_timing() = &t;
      hooked = true;
    }
  }
  hipEvent_t take() {
    if (!p->event_pool->empty()) {
      hipEvent_t e = p->event_pool->back();
      p->event_pool->pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
  }
  ~ProfileScope() {
    if (!p->profiling) return;
    if (hooked) launch_timing() = outer;
    const bool timed = t.start && t.stop && (!hooked || t.launches > 0);
    if (timed) {
      if (!hooked) (void)hipEventRecord(t.stop, s);
      p->pending[kclass].push_back({t.start, t.stop});
    } else {
      if (t.start) p->event_pool->push_back(t.start);
      if (t.stop) p->event_pool->push_back(t.stop);
    }
  }
};

int kernel_class_of(int axis, int mode) {
  if (axis == 2) return mode == MODE_FWD ? WFS_K_Z_FWD : mode == MODE_INV ? WFS_K_Z_INV : WFS_K_Z_IVF;
  if (axis == 1) return mode == MODE_FWD ? WFS_K_Y_FWD : WFS_K_Y_INV;
  return mode == MODE_FKI ? WFS_K_X_FKI : mode == MODE_FWD ? WFS_K_X_FWD : WFS_K_X_INV;
}

// one axis pass; axis 0 = x (strided, whole grid), 1 = y (strided per x slab), 2 = z (rows)
// sel picks the copy of the per-axis tables (the step index)
const void* table_copy(const wfs_plan* p, const void* base, int axis, int sel) {
  const size_t off = size_t(sel % p->ntab) * size_t(p->n[axis]) * p->elem;
  return static_cast<const char*>(base) + off;
}

int run_pass(wfs_plan* p, int axis, int mode, void* data, const void* pv, const void* const* ktab,
             double scale, hipStream_t s, int sel = 0) {
  PassParams pp;
  pp.tw = table_copy(p, p->tw[axis], axis, sel);
  pp.scale = scale;
  long outer = 1;
  int layout;
  // stacked packets simply add rows (z pass), x slabs (y pass) or outer slabs (x pass)
  if (axis == 2) {
    layout = LAYOUT_ROWS;
    pp.rows = long(p->batch) * p->n[0] * p->n[1];
  } else if (axis == 1) {
    layout = LAYOUT_COLS;
    outer = long(p->batch) * p->n[0];
    pp.inner = p->n[2];
  } else {
    layout = LAYOUT_COLS;
    outer = p->batch;
    pp.inner = long(p->n[1]) * p->n[2];
  }
  if (mode == MODE_FKI) {
    pp.kx = table_copy(p, ktab[0], 0, sel);
    pp.ky = table_copy(p, ktab[1], 1, sel);
    pp.kz = table_copy(p, ktab[2], 2, sel);
    pp.nz = p->n[2];
  }
  if (mode == MODE_IVF) {
    if (p->sep_set) {
      const size_t per = size_t(p->batch) * p->elem;
      const int c = sel % p->ntab;
      pp.spx = static_cast<const char*>(p->sep[0]) + size_t(c) * per * p->n[0];
      pp.spy = static_cast<const char*>(p->sep[1]) + size_t(c) * per * p->n[1];
      pp.spz = static_cast<const char*>(p->sep[2]) + size_t(c) * per * p->n[2];
      pp.snx = p->n[0];
      pp.sny = p->n[1];
    } else {
      pp.pv = pv;
    }
  }
  const int kc = kernel_class_of(axis, mode);
  ProfileScope scope(p, kc, s);
  return launch_fft_pass(p->precision, p->n[axis], layout, mode, data, pp, outer, p->variant[kc], s);
}

size_t grid_points(const wfs_plan* p) { return size_t(p->batch) * p->n[0] * p->n[1] * p->n[2]; }

bool slab_on(const wfs_plan* p) { return p->variant[WFS_K_YZ_SLAB] != 0; }

int alloc_table(wfs_plan* p, void** t, hipStream_t s) {
  if (*t) return 0;
  WFS_TRY(refuse_if_capturing(s, "the potential phase table"));
  hipError_t e = hipMalloc(t, grid_points(p) * p->elem);
  if (e != hipSuccess) { *t = nullptr; set_last_error("hipMalloc for the potential phase table failed"); return WFS_ERR_ALLOC; }
  return 0;
}

// the map of x planes of pv_lines that repeat the previous plane (kernel class WFS_K_PV_PLANES,
// variant 1): built with the lines table, so that graph capture finds it ready.  Under capture
// it is not built (its kernels would run only at replay, and an uncaptured call before that
// would read an unwritten map): the captured steps read every plane's own copy, and the next
// uncaptured call builds it.  Once the table itself has been written inside a captured graph
// (a potential set or a reorder under capture), every replay may change it behind the map, so
// from then on the plan reads every plane's own copy.
int share_planes(wfs_plan* p, hipStream_t s) {
  p->pv_planes_ok = false;
  if (p->variant[WFS_K_PV_PLANES] == 0 || p->pv_in_graph) return 0;
  if (capturing(s)) return 0;
  const long planes = long(p->batch) * p->n[0];
  if (!p->pv_planes) {
    if (hipMalloc(&p->pv_planes, size_t(2 * planes) * sizeof(int)) != hipSuccess) {
      p->pv_planes = nullptr;
      set_last_error("hipMalloc for the potential plane map failed");
      return WFS_ERR_ALLOC;
    }
  }
  WFS_TRY(launch_plane_dedup(p->pv_lines, long(p->n[1]) * p->n[2], p->n[0], p->batch, 0, planes,
                             p->pv_planes + planes, p->pv_planes, s));
  p->pv_planes_ok = true;
  return 0;
}

// the phase table in the order the current scheme reads (the other order is rebuilt from it
// on demand)
int set_potential(wfs_plan* p, const void* V, hipStream_t s) {
  const long count = long(grid_points(p));
  if (slab_on(p)) {
    WFS_TRY(alloc_table(p, &p->pv_lines, s));
    WFS_TRY(launch_potential_phase_lines(p->precision, count, static_cast<const float2*>(V), p->dt, p->pv_lines,
                                         slab_lines_order(p->n[1]), s));
    p->pv_lines_ok = true;
    p->pv_nat_ok = false;
    if (capturing(s)) p->pv_in_graph = true;
    WFS_TRY(share_planes(p, s));
  } else {
    WFS_TRY(alloc_table(p, &p->pv, s));
    WFS_TRY(launch_potential_phase(p->precision, count, static_cast<const float2*>(V), p->dt, p->pv, s));
    p->pv_nat_ok = true;
    p->pv_lines_ok = false;
  }
  p->pv_set = true;
  p->sep_set = false;
  return 0;
}

// x planes [x_begin, x_end) of V changed since the current table was built from V (one packet):
// only those planes of the table are rebuilt, in the order the steps read, and only the map
// flags of planes x_begin .. x_end (plane x_end is compared with its changed predecessor).
// The table's values are those set_potential writes: every element depends on its own voxel
// of V alone.  Without a current table of V the whole table is built.
int set_potential_planes(wfs_plan* p, const void* V, int x_begin, int x_end, hipStream_t s) {
  if (x_begin == x_end) return 0;
  const bool current = p->pv_set && !p->sep_set && (slab_on(p) ? p->pv_lines_ok : p->pv_nat_ok);
  if (!current) return set_potential(p, V, s);
  const long plane = long(p->n[1]) * p->n[2];
  const long off = long(x_begin) * plane, count = long(x_end - x_begin) * plane;
  const float2* v = static_cast<const float2*>(V) + off;
  if (slab_on(p)) {
    WFS_TRY(launch_potential_phase_lines(p->precision, count, v, p->dt, static_cast<char*>(p->pv_lines) + off * p->elem,
                                         slab_lines_order(p->n[1]), s));
    p->pv_nat_ok = false;
    if (capturing(s)) p->pv_in_graph = true;
    if (p->pv_planes_ok && !capturing(s)) {
      const long planes = p->n[0];
      const long last = x_end < p->n[0] ? x_end : p->n[0] - 1;
      WFS_TRY(launch_plane_dedup(p->pv_lines, plane, p->n[0], 1, x_begin, last - x_begin + 1, p->pv_planes + planes,
                                 p->pv_planes, s));
    } else {
      WFS_TRY(share_planes(p, s));
    }
  } else {
    WFS_TRY(launch_potential_phase(p->precision, count, v, p->dt, static_cast<char*>(p->pv) + off * p->elem, s));
    p->pv_lines_ok = false;
  }
  return 0;
}

// the table the next steps read: C order for the axis passes and the plane pass, lines order
// for the slab scheme.  A reorder issued under graph capture runs only at replay, so it does
// not mark the other order current: the next uncaptured call reorders again.
int current_table(wfs_plan* p, hipStream_t s, const void** out) {
  *out = nullptr;
  if (p->sep_set) return 0;
  const long count = long(grid_points(p));
  if (slab_on(p)) {
    if (!p->pv_lines_ok) {
      WFS_TRY(alloc_table(p, &p->pv_lines, s));
      WFS_TRY(launch_table_reorder(p->precision, count, p->pv, p->pv_lines, slab_lines_order(p->n[1]), true, s));
      p->pv_lines_ok = !capturing(s);
      p->pv_planes_ok = false;
      if (capturing(s)) p->pv_in_graph = true;
    }
    if (!p->pv_planes_ok && p->variant[WFS_K_PV_PLANES] != 0) WFS_TRY(share_planes(p, s));
    *out = p->pv_lines;
  } else {
    if (!p->pv_nat_ok) {
      WFS_TRY(alloc_table(p, &p->pv, s));
      WFS_TRY(launch_table_reorder(p->precision, count, p->pv_lines, p->pv, slab_lines_order(p->n[1]), false, s));
      p->pv_nat_ok = !capturing(s);
    }
    *out = p->pv;
  }
  return 0;
}

// per-axis phase tables exp(-i dt f_d) of a separable real potential (host, long double phase
// reduced mod 2 pi; dithered copies for complex64 plans)
int set_potential_axes(wfs_plan* p, const double* const f[3]) {
  const long double two_pi = 2.0L * 3.14159265358979323846264338327950288L;
  for (int a = 0; a < 3; ++a) {
    const long len = long(p->batch) * p->n[a];
    std::vector<double> re(len), im(len);
    for (long j = 0; j < len; ++j) {
      if (!std::isfinite(f[a][j])) { set_last_error("wfs_set_potential_axes: non-finite potential"); return WFS_ERR_INVALID; }
      long double th = -(long double)p->dt * (long double)f[a][j];
      th = fmodl(th, two_pi);
      re[j] = double(cosl(th));
      im[j] = double(sinl(th));
    }
    if (p->sep[a]) { (void)hipFree(p->sep[a]); p->sep[a] = nullptr; }
    WFS_TRY(upload_dithered(&p->sep[a], re, im, p->precision));
  }
  p->sep_set = true;
  p->pv_set = true;
  return 0;
}

// y inverse (table copy sel - 1), z inverse / potential / z forward and y forward (copy sel)
// of one step as one launch on whole x planes
int run_yz_plane(wfs_plan* p, void* data, const void* pv, hipStream_t s, int sel) {
  YzPlaneParams pp;
  pp.tw_yi = table_copy(p, p->tw[1], 1, sel - 1);
  pp.tw_yf = table_copy(p, p->tw[1], 1, sel);
  pp.tw_z = table_copy(p, p->tw[2], 2, sel);
  if (p->sep_set) {
    const size_t per = size_t(p->batch) * p->elem;
    const int c = sel % p->ntab;
    pp.spx = static_cast<const char*>(p->sep[0]) + size_t(c) * per * p->n[0];
    pp.spy = static_cast<const char*>(p->sep[1]) + size_t(c) * per * p->n[1];
    pp.spz = static_cast<const char*>(p->sep[2]) + size_t(c) * per * p->n[2];
  } else {
    pp.pv = pv;
  }
  pp.nx = p->n[0];
  pp.planes = long(p->batch) * p->n[0];
  ProfileScope scope(p, WFS_K_YZ_PLANE, s);
  return launch_yz_plane(p->precision, p->n[1], p->variant[WFS_K_YZ_PLANE], data, pp, s);
}

// the passes between two x passes: fused plane kernel or three axis passes
int run_yz(wfs_plan* p, void* data, const void* pv, hipStream_t s, int sel) {
  if (p->variant[WFS_K_YZ_PLANE] != 0) return run_yz_plane(p, data, pv, s, sel);
  WFS_TRY(run_pass(p, 1, MODE_INV, data, nullptr, nullptr, 1.0, s, sel - 1));
  WFS_TRY(run_pass(p, 2, MODE_IVF, data, pv, nullptr, 1.0, s, sel));
  return run_pass(p, 1, MODE_FWD, data, nullptr, nullptr, 1.0, s, sel);
}

// Slab scheme (ny = nz = 256): the y-z launch of one step on rows y = 4 a + b (mode
// LINES_FULL), or the entry (z, y forward) / exit (y, z inverse) launch of a call
int run_slab_lines(wfs_plan* p, int mode, void* data, const void* pv, hipStream_t s, int sel) {
  YzPlaneParams pp;
  const SlabRows g = slab_rows(p->n[1]);
  const int c = sel % p->ntab;
  pp.tw_yi = static_cast<const char*>(p->tw_rows[0]) + size_t(c) * g.uy * g.n1 * p->elem;
  pp.tw_yf = pp.tw_yi;
  pp.tw_z = static_cast<const char*>(p->tw_rows[1]) + size_t(c) * g.uz * g.n1 * p->elem;
  if (mode == LINES_FULL) {
    if (p->sep_set) {
      const size_t per = size_t(p->batch) * p->elem;
      const int c = sel % p->ntab;
      pp.spx = static_cast<const char*>(p->sep[0]) + size_t(c) * per * p->n[0];
      pp.spy = static_cast<const char*>(p->sep[1]) + size_t(c) * per * p->n[1];
      pp.spz = static_cast<const char*>(p->sep[2]) + size_t(c) * per * p->n[2];
    } else {
      pp.pv = pv;
      // the map describes the plan's own lines table (the autotuner steps scratch tables)
      if (pv == p->pv_lines && p->pv_planes_ok && p->variant[WFS_K_PV_PLANES] != 0) pp.pv_planes = p->pv_planes;
    }
  }
  pp.nx = p->n[0];
  pp.planes = long(p->batch) * p->n[0];
  ProfileScope scope(p, mode == LINES_FULL ? WFS_K_YZ_SLAB : WFS_K_SLAB_EDGE, s);
  return launch_yz_slab(p->n[1], mode, p->variant[WFS_K_YZ_SLAB], data, pp, s);
}

// Slab scheme: length-4 y stages around x forward -> K -> x inverse
int run_x_slab(wfs_plan* p, void* data, const void* const* ktab, hipStream_t s, int sel) {
  XSlabParams xp;
  xp.tw_x = table_copy(p, p->tw[0], 0, sel);
  xp.tw_y = table_copy(p, p->tw[1], 1, sel);
  xp.kx = table_copy(p, ktab[0], 0, sel);
  xp.ky = table_copy(p, ktab[1], 1, sel);
  xp.kz = table_copy(p, ktab[2], 2, sel);
  xp.packets = p->batch;
  ProfileScope scope(p, WFS_K_X_SLAB, s);
  return launch_x_slab(p->n[0], p->n[1], p->variant[WFS_K_X_SLAB], data, xp, s);
}

// Slab scheme, observing launches: LINES_OBSERVE runs the exit launch of one stretch (table copy
// sel_exit), the box moments of the position-space state and the entry launch of the next
// stretch (copy sel_entry) as one launch; LINES_ENTRY_OBSERVE the moments of the position-space
// input and the entry launch of a call.  The records are then summed into out[packet][box][7].
// The records (obs_lines) are allocated with the plan.
int run_slab_observe(wfs_plan* p, int mode, void* data, int sel_exit, int sel_entry, const wfs_box* boxes, int nb,
                     double* out, hipStream_t s) {
  const SlabRows g = slab_rows(p->n[1]);
  const long groups = long(p->batch) * p->n[0] * g.b;
  if (p->obs_lines_len < groups * kBoxesPerPass * 7) {
    set_last_error("wfs_step_observe: observing records missing");
    return WFS_ERR_INVALID;
  }
  YzPlaneParams pp;
  const size_t ysz = size_t(g.uy) * g.n1 * p->elem, zsz = size_t(g.uz) * g.n1 * p->elem;
  const char* ty = static_cast<const char*>(p->tw_rows[0]);
  const char* tz = static_cast<const char*>(p->tw_rows[1]);
  if (mode == LINES_OBSERVE) {
    pp.tw_yi = ty + size_t(sel_exit % p->ntab) * ysz;
    pp.tw_z = tz + size_t(sel_exit % p->ntab) * zsz;
    pp.tw_yf = ty + size_t(sel_entry % p->ntab) * ysz;
    pp.tw_zf = tz + size_t(sel_entry % p->ntab) * zsz;
  } else {
    pp.tw_yi = pp.tw_yf = ty + size_t(sel_entry % p->ntab) * ysz;
    pp.tw_z = tz + size_t(sel_entry % p->ntab) * zsz;
  }
  pp.nx = p->n[0];
  pp.planes = long(p->batch) * p->n[0];
  pp.mom.nb = nb;
  for (int b = 0; b < nb; ++b)
    for (int a = 0; a < 3; ++a) {
      pp.mom.bot[b][a] = boxes[b].bottom[a];
      pp.mom.num[b][a] = boxes[b].count[a];
    }
  for (int a = 0; a < 3; ++a) {
    pp.mom.n[a] = p->n[a];
    pp.mom.dx[a] = p->dx[a];
  }
  pp.mom.partials = p->obs_lines;
  {
    ProfileScope scope(p, WFS_K_SLAB_OBSERVE, s);
    WFS_TRY(launch_yz_slab(p->n[1], mode, 1, data, pp, s));
  }
  ProfileScope scope(p, WFS_K_OBSERVABLES, s);
  return launch_sum_partials(p->obs_lines, long(p->n[0]) * g.b, nb, p->batch, p->dx[0] * p->dx[1] * p->dx[2] / 27.0,
                             out, 7L * nb, s);
}

int observe(wfs_plan* p, const void* psi, const wfs_box* boxes, int nboxes, double* out, hipStream_t s);

int fourier_steps(wfs_plan* p, void* psi, const void* V, int nsteps, hipStream_t s) {
  if (V) WFS_TRY(set_potential(p, V, s));
  if (!p->pv_set) { set_last_error("wfs_step: no potential (pass V or call wfs_set_potential first)"); return WFS_ERR_INVALID; }
  const void* pv = nullptr;
  WFS_TRY(current_table(p, s, &pv));
  const void* const* kh = p->khalf;
  const void* const* kf = p->kfull;
  // every pass of the call's step k uses table copy b + k (mod ntab), the entry passes copy b:
  // the plan's running index b makes a sequence of short calls cycle the dithered copies as
  // one long call does
  const int b = p->step_index;
  p->step_index = int((long(b) + nsteps) % p->ntab);
  if (p->variant[WFS_K_YZ_SLAB] != 0) {
    // two launches per step; between them y holds the half-transformed Y[b][ka]
    WFS_TRY(run_slab_lines(p, LINES_ENTRY, psi, nullptr, s, b));
    WFS_TRY(run_x_slab(p, psi, kh, s, b));
    for (int step = 1; step <= nsteps; ++step) {
      WFS_TRY(run_slab_lines(p, LINES_FULL, psi, pv, s, b + step));
      WFS_TRY(run_x_slab(p, psi, step == nsteps ? kh : kf, s, b + step));
    }
    return run_slab_lines(p, LINES_EXIT, psi, nullptr, s, b + nsteps);
  }
  WFS_TRY(run_pass(p, 2, MODE_FWD, psi, nullptr, nullptr, 1.0, s, b));
  WFS_TRY(run_pass(p, 1, MODE_FWD, psi, nullptr, nullptr, 1.0, s, b));
  WFS_TRY(run_pass(p, 0, MODE_FKI, psi, nullptr, kh, 1.0, s, b));
  for (int step = 1; step <= nsteps; ++step) {
    WFS_TRY(run_yz(p, psi, pv, s, b + step));
    WFS_TRY(run_pass(p, 0, MODE_FKI, psi, nullptr, step == nsteps ? kh : kf, 1.0, s, b + step));
  }
  WFS_TRY(run_pass(p, 1, MODE_INV, psi, nullptr, nullptr, 1.0, s, b + nsteps));
  WFS_TRY(run_pass(p, 2, MODE_INV, psi, nullptr, nullptr, 1.0, s, b + nsteps));
  return 0;
}

// nsteps steps with the box observables of the state before steps 0, every, 2 every, ...
// (< nsteps) written to out[o][packet][box][7]: the same launches and results as alternating
// wfs_observables and wfs_step(every) calls.  The slab scheme runs the exit of one call, the
// observation and the entry of the next as one launch (run_slab_observe).
int steps_observed(wfs_plan* p, void* psi, const void* V, int nsteps, int every, const wfs_box* boxes, int nb,
                   double* out, hipStream_t s, wfs_between_steps between = nullptr, void* ctx = nullptr);

int power_series_steps(wfs_plan* p, void* psi, const void* V, int nsteps, hipStream_t s) {
  if (!V) { set_last_error("wfs_step: the power series needs V"); return WFS_ERR_INVALID; }
  WFS_TRY(ensure_scratch(p, s));
  const double inv_2m = 1.0 / (2.0 * p->mass);
  const double inv_dx2 = 1.0 / (p->dx[0] * p->dx[0]);   // x spacing on every axis (power_series_operator.cu:58)
  const size_t bytes = grid_points(p) * p->elem;
  // variant 1 (default): stages in fused pairs; variant 0: one launch per stage.  Both give
  // the same bits.  A pair reads psi around its tile, so it never writes psi in place; the
  // result of the last pair is copied back, a last single stage writes psi directly.
  const bool pairs = p->variant[WFS_K_HORNER] != 0;
  for (int step = 0; step < nsteps; ++step) {
    const void* u_in = psi;
    int which = 0;
    int nn = p->order;
    while (nn >= 1) {
      ProfileScope scope(p, WFS_K_HORNER, s);
      if (pairs && nn >= 2) {
        void* out = p->scratch[which];
        WFS_TRY(launch_horner_pair(p->precision, p->batch, p->n[0], p->n[1], p->n[2], u_in, out, psi,
                                   static_cast<const float2*>(V), p->dt / double(nn), p->dt / double(nn - 1),
                                   inv_2m, inv_dx2, p->variant[WFS_K_HORNER], s));
        u_in = out;
        nn -= 2;
      } else {
        void* out = (nn == 1 && u_in != psi) ? psi : p->scratch[which];
        WFS_TRY(launch_horner_stage(p->precision, p->batch, p->n[0], p->n[1], p->n[2], u_in, out, psi,
                                    static_cast<const float2*>(V), p->dt / double(nn), inv_2m, inv_dx2, 0, s));
        u_in = out;
        nn -= 1;
      }
      which ^= 1;
    }
    if (u_in != psi) WFS_HIP_CHECK(hipMemcpyAsync(psi, u_in, bytes, hipMemcpyDeviceToDevice, s));
  }
  return 0;
}

}  // namespace
}  // namespace wfs

using namespace wfs;

extern "C" const char* wfs_version(void) { return "wfsim-gfx950 0.1.0"; }

extern "C" const char* wfs_last_error(void) { return g_last_error; }

extern "C" int wfs_plan_destroy(wfs_plan* p) {
  if (!p) return 0;
  for (int a = 0; a < 3; ++a) {
    if (p->tw[a]) (void)hipFree(p->tw[a]);
    if (p->kfull[a]) (void)hipFree(p->kfull[a]);
    if (p->khalf[a]) (void)hipFree(p->khalf[a]);
    if (p->sep[a]) (void)hipFree(p->sep[a]);
  }
  for (int i = 0; i < 2; ++i)
    if (p->scratch[i]) (void)hipFree(p->scratch[i]);
  if (p->obs_work) (void)hipFree(p->obs_work);
  if (p->plane_work) (void)hipFree(p->plane_work);
  if (p->plane_out) (void)hipFree(p->plane_out);
  if (p->pv) (void)hipFree(p->pv);
  if (p->pv_lines) (void)hipFree(p->pv_lines);
  if (p->pv_planes) (void)hipFree(p->pv_planes);
  if (p->obs_lines) (void)hipFree(p->obs_lines);
  for (int k = 0; k < 2; ++k)
    if (p->tw_rows[k]) (void)hipFree(p->tw_rows[k]);
  if (p->pending) {
    for (int k = 0; k < WFS_K_COUNT; ++k)
      for (auto& pr : p->pending[k]) { (void)hipEventDestroy(pr.first); (void)hipEventDestroy(pr.second); }
    delete[] p->pending;
  }
  if (p->event_pool) {
    for (auto e : *p->event_pool) (void)hipEventDestroy(e);
    delete p->event_pool;
  }
  delete p;
  return 0;
}

extern "C" int wfs_plan_create(const int n[3], const double dx[3], double dt, double mass, int precision,
                               int method, int order, wfs_plan** out) {
  return wfs_plan_create_batched(n, dx, dt, mass, precision, method, order, 1, out);
}

extern "C" int wfs_plan_create_batched(const int n[3], const double dx[3], double dt, double mass, int precision,
                                       int method, int order, int batch, wfs_plan** out) {
  if (!out || !n || !dx) { set_last_error("wfs_plan_create: null argument"); return WFS_ERR_INVALID; }
  if (batch < 1) { set_last_error("wfs_plan_create: batch must be >= 1"); return WFS_ERR_INVALID; }
  *out = nullptr;
  for (int a = 0; a < 3; ++a) {
    if (!fft_length_supported(n[a])) {
      set_last_error("wfs_plan_create: axis %d length %d not supported (8,16,32,64,96,128,192,256,512)", a, n[a]);
      return WFS_ERR_UNSUPPORTED;
    }
    if (!(dx[a] > 0.0)) { set_last_error("wfs_plan_create: dx[%d] must be positive", a); return WFS_ERR_INVALID; }
  }
  if (precision != WFS_C64 && precision != WFS_C128) { set_last_error("wfs_plan_create: bad precision %d", precision); return WFS_ERR_INVALID; }
  if (method != WFS_METHOD_FOURIER && method != WFS_METHOD_POWER_SERIES) { set_last_error("wfs_plan_create: bad method %d", method); return WFS_ERR_INVALID; }
  if (method == WFS_METHOD_POWER_SERIES && order < 1) { set_last_error("wfs_plan_create: order must be >= 1"); return WFS_ERR_INVALID; }
  if (!(mass > 0.0)) { set_last_error("wfs_plan_create: mass must be positive"); return WFS_ERR_INVALID; }

  wfs_plan* p = new wfs_plan();
  std::memset(p, 0, sizeof(*p));
  for (int a = 0; a < 3; ++a) { p->n[a] = n[a]; p->dx[a] = dx[a]; }
  p->batch = batch;
  p->dt = dt; p->mass = mass; p->precision = precision; p->method = method; p->order = order;
  p->elem = precision == WFS_C128 ? 16 : 8;
  p->ntab = precision == WFS_C128 ? 1 : kDither;
  for (int k = 0; k < WFS_K_COUNT; ++k) p->variant[k] = 1;
  // fused stage pairs: 8 x 128 tiles where the z axis fills them (measured fastest at 256^3),
  // else 16 x 64
  p->variant[WFS_K_HORNER] = p->n[2] >= 256 ? 5 : p->n[2] >= 128 ? 4 : 1;
  // fused plane pass where the plane fits in LDS; at 128^2 two rounds (two blocks per CU) once
  // there are enough planes to fill the CUs twice over (measured: 1024 planes 13% faster, 128
  // planes 11% slower)
  p->variant[WFS_K_YZ_PLANE] =
      (yz_plane_supported(precision, n[1], n[2], 2) && n[1] == 128 && long(batch) * n[0] >= 512) ? 2
      : yz_plane_supported(precision, n[1], n[2], 1) ? 1 : 0;
  // slab scheme at ny = nz = 256, 192 or 128 (complex64).  At 256 the persistent y-z launch
  // and 8-z x blocks (measured fastest at 256^3: 90 and 66 us against 115 and 79 for the other
  // forms); at 192 one block per row group (two 72 KiB blocks per CU) and 16-z x blocks
  // (192^3 x 15: 662 and 460 us; 8-z x blocks 514 us); at 128 only while the x planes are too
  // few to fill the CUs with plane-pass blocks, fewer than 256 (128^3: 29 against 34 us per
  // step; from 8 packets on the plane pass is as fast)
  const bool slab = slab_supported(precision, n) && (n[1] != 128 || long(batch) * n[0] < 256);
  p->variant[WFS_K_YZ_SLAB] = slab ? (n[1] == 192 ? 1 : 2) : 0;
  if (slab) p->variant[WFS_K_X_SLAB] = n[1] == 192 ? 1 : 2;
  p->pending = new std::vector<std::pair<hipEvent_t, hipEvent_t>>[WFS_K_COUNT];
  p->event_pool = new std::vector<hipEvent_t>();
  int rc = 0;
  for (int a = 0; a < 3 && rc == 0; ++a) {
    std::vector<double> re(n[a]), im(n[a]);
    for (long j = 0; j < n[a]; ++j) {
      double c, s;
      unit_root(j, n[a], c, s);
      re[j] = c; im[j] = -s;   // forward twiddle exp(-2 pi i j / N)
    }
    rc = upload_dithered(&p->tw[a], re, im, precision);
    if (rc) break;
    if (a > 0 && slab_supported(precision, n)) {
      const SlabRows g = slab_rows(n[1]);
      rc = a == 1 ? upload_rows_table(&p->tw_rows[0], re, im, g.b, g.uy, g.n1)
                  : upload_rows_table(&p->tw_rows[1], re, im, 1, g.uz, g.n1);
      if (rc) break;
    }
    kinetic_table(n[a], dx[a], dt, mass, 2.0, re, im);
    rc = upload_dithered(&p->kfull[a], re, im, precision);
    if (rc) break;
    kinetic_table(n[a], dx[a], dt, mass, 4.0, re, im);
    rc = upload_dithered(&p->khalf[a], re, im, precision);
  }
  if (rc == 0) {
    wfs_box whole{{0, 0, 0}, {n[0], n[1], n[2]}};
    p->obs_work_len = box_moments_partials_needed(whole) * batch;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&p->obs_work), size_t(p->obs_work_len) * sizeof(double));
    if (e != hipSuccess) { set_last_error("hipMalloc for observable workspace failed"); rc = WFS_ERR_ALLOC; }
  }
  if (rc == 0) {
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&p->plane_out), 7 * sizeof(double));
    if (e != hipSuccess) { set_last_error("hipMalloc for the plane-current result failed"); rc = WFS_ERR_ALLOC; }
  }
  if (rc == 0 && slab_supported(precision, n)) {
    // box-moment records of the slab scheme's observing launch: one per row group and box
    const long need = long(batch) * n[0] * slab_rows(n[1]).b * kBoxesPerPass * 7;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&p->obs_lines), size_t(need) * sizeof(double));
    if (e != hipSuccess) { set_last_error("hipMalloc for observing records failed"); rc = WFS_ERR_ALLOC; }
    else p->obs_lines_len = need;
  }
  if (rc) { wfs_plan_destroy(p); return rc; }
  *out = p;
  return 0;
}

extern "C" int wfs_set_potential(wfs_plan* p, const void* V, void* stream) {
  if (!p || !V) { set_last_error("wfs_set_potential: null argument"); return WFS_ERR_INVALID; }
  if (p->method != WFS_METHOD_FOURIER) return 0;   // the power series reads V directly
  return set_potential(p, V, static_cast<hipStream_t>(stream));
}

extern "C" int wfs_set_potential_planes(wfs_plan* p, const void* V, int x_begin, int x_end, void* stream) {
  if (!p || !V) { set_last_error("wfs_set_potential_planes: null argument"); return WFS_ERR_INVALID; }
  if (x_begin < 0 || x_begin > x_end || x_end > p->n[0]) {
    set_last_error("wfs_set_potential_planes: planes [%d, %d) outside 0 .. %d", x_begin, x_end, p->n[0]);
    return WFS_ERR_INVALID;
  }
  if (p->batch != 1) { set_last_error("wfs_set_potential_planes: batched plans are not supported"); return WFS_ERR_UNSUPPORTED; }
  if (p->method != WFS_METHOD_FOURIER) return 0;   // the power series reads V directly
  return set_potential_planes(p, V, x_begin, x_end, static_cast<hipStream_t>(stream));
}

extern "C" int wfs_set_potential_axes(wfs_plan* p, const double* fx, const double* fy, const double* fz) {
  if (!p || !fx || !fy || !fz) { set_last_error("wfs_set_potential_axes: null argument"); return WFS_ERR_INVALID; }
  if (p->method != WFS_METHOD_FOURIER) { set_last_error("wfs_set_potential_axes: split-step plans only"); return WFS_ERR_UNSUPPORTED; }
  const double* const f[3] = {fx, fy, fz};
  return set_potential_axes(p, f);
}

extern "C" int wfs_step(wfs_plan* p, void* psi, const void* V, int nsteps, void* stream) {
  if (!p || !psi) { set_last_error("wfs_step: null argument"); return WFS_ERR_INVALID; }
  if (nsteps < 0) { set_last_error("wfs_step: negative step count"); return WFS_ERR_INVALID; }
  if (nsteps == 0) return 0;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (p->method == WFS_METHOD_FOURIER) return fourier_steps(p, psi, V, nsteps, s);
  return power_series_steps(p, psi, V, nsteps, s);
}

extern "C" int wfs_fft3d(wfs_plan* p, void* data, int direction, void* stream) {
  if (!p || !data || (direction != -1 && direction != 1)) { set_last_error("wfs_fft3d: bad argument"); return WFS_ERR_INVALID; }
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int mode = direction < 0 ? MODE_FWD : MODE_INV;
  for (int axis = 2; axis >= 0; --axis) {
    const int rc = run_pass(p, axis, mode, data, nullptr, nullptr, 1.0 / std::sqrt(double(p->n[axis])), s);
    if (rc) return rc;
  }
  return 0;
}

namespace wfs {
namespace {
int check_boxes(const wfs_plan* p, const wfs_box* boxes, int nboxes) {
  for (int b = 0; b < nboxes; ++b) {
    const wfs_box& bx = boxes[b];
    for (int a = 0; a < 3; ++a) {
      if (bx.bottom[a] < 0 || bx.count[a] < 1 || long(bx.bottom[a]) + bx.count[a] > p->n[a]) {
        set_last_error("wfs_observables: box %d axis %d [%d, %d) outside the grid of %d voxels", b, a,
                       bx.bottom[a], bx.bottom[a] + bx.count[a], p->n[a]);
        return WFS_ERR_INVALID;
      }
    }
  }
  return 0;
}

int observe(wfs_plan* p, const void* psi, const wfs_box* boxes, int nboxes, double* out, hipStream_t s) {
  WFS_TRY(check_boxes(p, boxes, nboxes));
  // up to kBoxesPerPass boxes per pass over their bounding box
  for (int b0 = 0; b0 < nboxes; b0 += kBoxesPerPass) {
    const int nb = std::min(kBoxesPerPass, nboxes - b0);
    ProfileScope scope(p, WFS_K_OBSERVABLES, s);
    const int rc = launch_box_moments(p->precision, p->batch, p->n[0], p->n[1], p->n[2], p->dx, psi, boxes + b0,
                                      nb, p->obs_work, p->obs_work_len, out + 7 * b0, 7L * nboxes, s);
    if (rc) return rc;
  }
  return 0;
}

// the caller's between-steps callback after step `step` (its kernels queued, the next step's not)
static int call_between(wfs_between_steps between, void* ctx, int step) {
  if (!between) return 0;
  const int rc = between(ctx, step);
  if (rc) {
    set_last_error("wfs_step_observe_between: the callback returned %d after step %d", rc, step);
    return WFS_ERR_INVALID;
  }
  return 0;
}

int steps_observed(wfs_plan* p, void* psi, const void* V, int nsteps, int every, const wfs_box* boxes, int nb,
                   double* out, hipStream_t s, wfs_between_steps between, void* ctx) {
  const long per_obs = long(p->batch) * nb * 7;
  const bool fused = p->method == WFS_METHOD_FOURIER && p->variant[WFS_K_YZ_SLAB] != 0 && nb <= kBoxesPerPass &&
                     nb > 0;
  if (!fused) {
    // with a callback, one step per call, so that it runs between every two steps
    const int chunk = between ? 1 : every;
    for (int k = 0; k < nsteps; k += chunk) {
      if (nb > 0 && k % every == 0) WFS_TRY(observe(p, psi, boxes, nb, out + (k / every) * per_obs, s));
      const int m = std::min(chunk, nsteps - k);
      WFS_TRY(p->method == WFS_METHOD_FOURIER ? fourier_steps(p, psi, k == 0 ? V : nullptr, m, s)
                                              : power_series_steps(p, psi, V, m, s));
      WFS_TRY(call_between(between, ctx, k));
    }
    return 0;
  }
  if (V) WFS_TRY(set_potential(p, V, s));
  if (!p->pv_set) { set_last_error("wfs_step_observe: no potential (pass V or call wfs_set_potential first)"); return WFS_ERR_INVALID; }
  WFS_TRY(check_boxes(p, boxes, nb));
  const void* pv = nullptr;
  WFS_TRY(current_table(p, s, &pv));
  // table copies as in the alternating calls: the stretch that starts at step k uses b + k
  const int b = p->step_index;
  p->step_index = int((long(b) + nsteps) % p->ntab);
  // the first observation rides on the entry launch
  WFS_TRY(run_slab_observe(p, LINES_ENTRY_OBSERVE, psi, b, b, boxes, nb, out, s));
  WFS_TRY(run_x_slab(p, psi, p->khalf, s, b));
  for (int k = 0; k < nsteps; k += every) {
    const int m = std::min(every, nsteps - k);
    for (int step = 1; step <= m; ++step) {
      WFS_TRY(run_slab_lines(p, LINES_FULL, psi, pv, s, b + k + step));
      WFS_TRY(run_x_slab(p, psi, step == m ? p->khalf : p->kfull, s, b + k + step));
      // the table is read next by step k + step's y-z launch: what the callback queues on the
      // stream to rewrite it runs after this step's and before that one
      WFS_TRY(call_between(between, ctx, k + step - 1));
    }
    if (k + m < nsteps) {
      WFS_TRY(run_slab_observe(p, LINES_OBSERVE, psi, b + k + m, b + k + m, boxes, nb,
                               out + ((k + m) / every) * per_obs, s));
      WFS_TRY(run_x_slab(p, psi, p->khalf, s, b + k + m));
    } else {
      WFS_TRY(run_slab_lines(p, LINES_EXIT, psi, nullptr, s, b + k + m));
    }
  }
  return 0;
}
}  // namespace
}  // namespace wfs

extern "C" int wfs_observables(wfs_plan* p, const void* psi, const wfs_box* boxes, int nboxes, double* out,
                               void* stream) {
  if (!p || !psi || (nboxes > 0 && (!boxes || !out)) || nboxes < 0) { set_last_error("wfs_observables: bad argument"); return WFS_ERR_INVALID; }
  return observe(p, psi, boxes, nboxes, out, static_cast<hipStream_t>(stream));
}

extern "C" int wfs_step_observe(wfs_plan* p, void* psi, const void* V, int nsteps, int observe_every,
                                const wfs_box* boxes, int nboxes, double* out, void* stream) {
  if (!p || !psi || nsteps < 0 || observe_every < 1 || nboxes < 0 || (nboxes > 0 && (!boxes || !out))) {
    set_last_error("wfs_step_observe: bad argument");
    return WFS_ERR_INVALID;
  }
  if (nsteps == 0) return 0;
  if (p->method == WFS_METHOD_POWER_SERIES && !V) { set_last_error("wfs_step_observe: the power series needs V"); return WFS_ERR_INVALID; }
  return steps_observed(p, psi, V, nsteps, observe_every, boxes, nboxes, out, static_cast<hipStream_t>(stream));
}

extern "C" int wfs_step_observe_between(wfs_plan* p, void* psi, const void* V, int nsteps, int observe_every,
                                        const wfs_box* boxes, int nboxes, double* out, wfs_between_steps between,
                                        void* ctx, void* stream) {
  if (!p || !psi || nsteps < 0 || observe_every < 1 || nboxes < 0 || (nboxes > 0 && (!boxes || !out))) {
    set_last_error("wfs_step_observe_between: bad argument");
    return WFS_ERR_INVALID;
  }
  if (nsteps == 0) return 0;
  if (p->method == WFS_METHOD_POWER_SERIES && !V) {
    set_last_error("wfs_step_observe_between: the power series needs V");
    return WFS_ERR_INVALID;
  }
  return steps_observed(p, psi, V, nsteps, observe_every, boxes, nboxes, out, static_cast<hipStream_t>(stream),
                        between, ctx);
}

extern "C" int wfs_plan_set_step_index(wfs_plan* p, long index) {
  if (!p || index < 0) { set_last_error("wfs_plan_set_step_index: bad argument"); return WFS_ERR_INVALID; }
  p->step_index = int(index % p->ntab);
  return 0;
}

extern "C" long wfs_plan_get_step_index(wfs_plan* p) {
  if (!p) return -1;
  return p->step_index;
}

extern "C" int wfs_plan_set_profiling(wfs_plan* p, int enable) {
  if (!p) { set_last_error("wfs_plan_set_profiling: null plan"); return WFS_ERR_INVALID; }
  p->profiling = enable != 0;
  return 0;
}

extern "C" int wfs_plan_profile(wfs_plan* p, int kc, int reset, long* launches, double* total_ms) {
  if (!p || kc < 0 || kc >= WFS_K_COUNT) { set_last_error("wfs_plan_profile: bad argument"); return WFS_ERR_INVALID; }
  auto& v = p->pending[kc];
  for (auto& pr : v) {
    WFS_HIP_CHECK(hipEventSynchronize(pr.second));
    float ms = 0.0f;
    WFS_HIP_CHECK(hipEventElapsedTime(&ms, pr.first, pr.second));
    p->total_ms[kc] += double(ms);
    p->launches[kc] += 1;
    p->event_pool->push_back(pr.first);
    p->event_pool->push_back(pr.second);
  }
  v.clear();
  if (launches) *launches = p->launches[kc];
  if (total_ms) *total_ms = p->total_ms[kc];
  if (reset) { p->launches[kc] = 0; p->total_ms[kc] = 0.0; }
  return 0;
}

extern "C" int wfs_plan_set_variant(wfs_plan* p, int kc, int variant) {
  const int vmax = kc == WFS_K_HORNER ? 5 : kc == WFS_K_PV_PLANES ? 1 : 2;
  if (!p || kc < 0 || kc >= WFS_K_COUNT || variant < 0 || variant > vmax) { set_last_error("wfs_plan_set_variant: bad argument"); return WFS_ERR_INVALID; }
  if (kc == WFS_K_YZ_PLANE && variant > 0 && !yz_plane_supported(p->precision, p->n[1], p->n[2], variant)) {
    set_last_error("wfs_plan_set_variant: fused plane pass variant %d unsupported for %d x %d planes", variant, p->n[1], p->n[2]);
    return WFS_ERR_UNSUPPORTED;
  }
  if (kc == WFS_K_YZ_SLAB && variant > 0 && !slab_supported(p->precision, p->n)) {
    set_last_error("wfs_plan_set_variant: the slab scheme needs complex64, ny = nz in {128, 192, 256} and nx <= 256");
    return WFS_ERR_UNSUPPORTED;
  }
  p->variant[kc] = variant;
  return 0;
}

extern "C" int wfs_plan_get_variant(wfs_plan* p, int kc) {
  if (!p || kc < 0 || kc >= WFS_K_COUNT) return -1;
  return p->variant[kc];
}

// Time every tile variant of each split-step pass on scratch data and keep
// the fastest.  Variants only change how many columns a block owns, so the
// arithmetic per point, and hence the result, is identical for all of them.
extern "C" int wfs_plan_autotune(wfs_plan* p, void* stream) {
  if (!p) { set_last_error("wfs_plan_autotune: null plan"); return WFS_ERR_INVALID; }
  if (p->method != WFS_METHOD_FOURIER) return 0;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const size_t bytes = grid_points(p) * p->elem;
  void* buf = nullptr;
  void* tab = nullptr;
  WFS_HIP_CHECK(hipMalloc(&buf, bytes));
  hipError_t e = hipMalloc(&tab, bytes);
  if (e != hipSuccess) { (void)hipFree(buf); set_last_error("wfs_plan_autotune: scratch allocation failed"); return WFS_ERR_ALLOC; }
  (void)hipMemsetAsync(buf, 0, bytes, s);
  (void)hipMemsetAsync(tab, 0, bytes, s);
  hipEvent_t a, b;
  (void)hipEventCreate(&a);
  (void)hipEventCreate(&b);
  const bool prof = p->profiling;
  p->profiling = false;
  struct Kind { int axis, mode; };
  const Kind kinds[] = {{2, MODE_FWD}, {2, MODE_INV}, {2, MODE_IVF}, {1, MODE_FWD}, {1, MODE_INV}, {0, MODE_FKI}};
  int rc = 0;
  for (const Kind& k : kinds) {
    const int kc = kernel_class_of(k.axis, k.mode);
    const int keep = p->variant[kc];
    float best = 1e30f;
    int best_v = keep;
    for (int v = 0; v < 3 && rc == 0; ++v) {
      p->variant[kc] = v;
      rc = run_pass(p, k.axis, k.mode, buf, tab, p->kfull, 1.0, s);   // warm-up
      (void)hipEventRecord(a, s);
      for (int r = 0; r < 3 && rc == 0; ++r) rc = run_pass(p, k.axis, k.mode, buf, tab, p->kfull, 1.0, s);
      (void)hipEventRecord(b, s);
      (void)hipEventSynchronize(b);
      float ms = 0.0f;
      (void)hipEventElapsedTime(&ms, a, b);
      if (rc == 0 && ms < best) { best = ms; best_v = v; }
    }
    p->variant[kc] = rc == 0 ? best_v : keep;
    if (rc) break;
  }
  // fused plane pass against the three (tuned) axis passes it replaces
  bool any_plane = false;
  for (int v = 1; v < 3; ++v) any_plane = any_plane || yz_plane_supported(p->precision, p->n[1], p->n[2], v);
  if (rc == 0 && any_plane) {
    const int keep = p->variant[WFS_K_YZ_PLANE];
    float best = 1e30f;
    int best_v = keep;
    for (int v = 0; v < 3 && rc == 0; ++v) {
      if (v > 0 && !yz_plane_supported(p->precision, p->n[1], p->n[2], v)) continue;
      p->variant[WFS_K_YZ_PLANE] = v;
      rc = run_yz(p, buf, tab, s, 1);
      (void)hipEventRecord(a, s);
      for (int r = 0; r < 3 && rc == 0; ++r) rc = run_yz(p, buf, tab, s, 1);
      (void)hipEventRecord(b, s);
      (void)hipEventSynchronize(b);
      float ms = 0.0f;
      (void)hipEventElapsedTime(&ms, a, b);
      if (rc == 0 && ms < best) { best = ms; best_v = v; }
    }
    p->variant[WFS_K_YZ_PLANE] = rc == 0 ? best_v : keep;
  }
  // slab scheme (two launches per step) against the tuned four axis passes: first the x
  // launch's block width, then one step of each y-z form and of the passes
  if (rc == 0 && slab_supported(p->precision, p->n)) {
    {
      const int keep = p->variant[WFS_K_X_SLAB];
      float best = 1e30f;
      int best_v = keep;
      for (int v = 1; v < 3 && rc == 0; ++v) {
        p->variant[WFS_K_X_SLAB] = v;
        rc = run_x_slab(p, buf, p->kfull, s, 1);
        (void)hipEventRecord(a, s);
        for (int r = 0; r < 3 && rc == 0; ++r) rc = run_x_slab(p, buf, p->kfull, s, 1);
        (void)hipEventRecord(b, s);
        (void)hipEventSynchronize(b);
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, a, b);
        if (rc == 0 && ms < best) { best = ms; best_v = v; }
      }
      p->variant[WFS_K_X_SLAB] = rc == 0 ? best_v : keep;
    }
    const int keep = p->variant[WFS_K_YZ_SLAB];
    float best = 1e30f;
    int best_v = keep;
    for (int v = 0; v < 3 && rc == 0; ++v) {
      p->variant[WFS_K_YZ_SLAB] = v;
      auto one_step = [&]() -> int {
        if (v) {
          WFS_TRY(run_slab_lines(p, LINES_FULL, buf, tab, s, 1));
          return run_x_slab(p, buf, p->kfull, s, 1);
        }
        WFS_TRY(run_yz(p, buf, tab, s, 1));
        return run_pass(p, 0, MODE_FKI, buf, nullptr, p->kfull, 1.0, s, 1);
      };
      rc = one_step();
      (void)hipEventRecord(a, s);
      for (int r = 0; r < 3 && rc == 0; ++r) rc = one_step();
      (void)hipEventRecord(b, s);
      (void)hipEventSynchronize(b);
      float ms = 0.0f;
      (void)hipEventElapsedTime(&ms, a, b);
      if (rc == 0 && ms < best) { best = ms; best_v = v; }
    }
    p->variant[WFS_K_YZ_SLAB] = rc == 0 ? best_v : keep;
  }
  p->profiling = prof;
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  (void)hipStreamSynchronize(s);
  (void)hipFree(buf);
  (void)hipFree(tab);
  return rc;
}

// ---------------------------------------------------------------------------
// hipGraph capture of a stream-ordered sequence of library calls
// ---------------------------------------------------------------------------
struct wfs_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

extern "C" int wfs_graph_begin(void* stream, wfs_graph** out) {
  if (!out || !stream) { set_last_error("wfs_graph_begin: needs a non-null stream"); return WFS_ERR_INVALID; }
  *out = nullptr;
  WFS_HIP_CHECK(hipStreamBeginCapture(static_cast<hipStream_t>(stream), hipStreamCaptureModeThreadLocal));
  *out = new wfs_graph();
  return 0;
}

extern "C" int wfs_graph_end(wfs_graph* g, void* stream) {
  if (!g || !stream) { set_last_error("wfs_graph_end: bad argument"); return WFS_ERR_INVALID; }
  WFS_HIP_CHECK(hipStreamEndCapture(static_cast<hipStream_t>(stream), &g->graph));
  WFS_HIP_CHECK(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
  return 0;
}

extern "C" int wfs_graph_launch(wfs_graph* g, void* stream) {
  if (!g || !g->exec) { set_last_error("wfs_graph_launch: graph not instantiated"); return WFS_ERR_INVALID; }
  WFS_HIP_CHECK(hipGraphLaunch(g->exec, static_cast<hipStream_t>(stream)));
  return 0;
}

extern "C" long wfs_graph_node_count(wfs_graph* g) {
  if (!g || !g->graph) return -1;
  size_t n = 0;
  if (hipGraphGetNodes(g->graph, nullptr, &n) != hipSuccess) return -1;
  return long(n);
}

extern "C" int wfs_graph_destroy(wfs_graph* g) {
  if (!g) return 0;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  delete g;
  return 0;
}

extern "C" int wfs_plane_current(wfs_plan* p, const void* psi, const double center[3], const double normal[3],
                                 const double size[2], const int count[2], void* density, double* out,
                                 void* stream) {
  if (!p || !psi || !center || !normal || !size || !count || !out || count[0] < 2 || count[1] < 2) {
    set_last_error("wfs_plane_current: bad argument");
    return WFS_ERR_INVALID;
  }
  if (p->batch != 1) { set_last_error("wfs_plane_current: batched plans are not supported"); return WFS_ERR_UNSUPPORTED; }
  hipStream_t s = static_cast<hipStream_t>(stream);
  const long need = plane_current_partials_needed(count[0], count[1]);
  if (need > p->plane_work_len) {
    // grown on demand: the pixel count is known only here
    WFS_TRY(refuse_if_capturing(s, "wfs_plane_current"));
    if (p->plane_work) { WFS_HIP_CHECK(hipStreamSynchronize(s)); (void)hipFree(p->plane_work); p->plane_work = nullptr; }
    WFS_HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&p->plane_work), size_t(need) * sizeof(double)));
    p->plane_work_len = need;
  }
  ProfileScope scope(p, WFS_K_OBSERVABLES, s);
  WFS_TRY(launch_plane_current(p->precision, p->n[0], p->n[1], p->n[2], p->dx, psi, center, normal, size, count,
                               p->mass, density, p->plane_work, p->plane_work_len, p->plane_out, s));
  WFS_HIP_CHECK(hipMemcpyAsync(out, p->plane_out, sizeof(double), hipMemcpyDeviceToDevice, s));
  return 0;
}
