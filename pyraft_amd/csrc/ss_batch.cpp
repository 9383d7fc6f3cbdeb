// ss_batch.cpp -- backprojection kernel variants and their autotuner, and batches
// of slices with different geometry run as persistent launches, optionally
// captured once into a HIP graph and replayed.
//
// A batch keeps one plan per distinct geometry, the packed-buffer offsets of every
// slice, and per launch kind (gather, transpose, Radon / fixup / backprojection over
// one angle range) a device table of entries (slice, angle range, first item).  The
// kernels walk the items of all slices with a fixed grid of resident blocks.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "ss_internal.h"

// ------------------------------------------------------------------ variants

extern "C" int raft_ss_bp_variant_shape(int v, int* rows, int* stage) {
  if (v < 0 || v >= SS_BP_NVARIANTS)
    return ss_set_error(RAFT_SS_EINVAL, "backprojection variant %d outside [0, %d)", v, SS_BP_NVARIANTS);
  if (rows) *rows = kSSBpRows[v];
  if (stage) *stage = kSSBpStage[v];
  return 0;
}

static int check_variant(const raft_ss_plan* p, int v) {
  if (v < 0 || v >= SS_BP_NVARIANTS)
    return ss_set_error(RAFT_SS_EINVAL, "backprojection variant %d outside [0, %d)", v, SS_BP_NVARIANTS);
  if (p->bp_angles_v[v] <= 0)
    return ss_set_error(RAFT_SS_EINVAL, "variant %d cannot stage this geometry (window %d > %d pairs)", v,
                        p->bp_window_v[v], kSSBpStage[v]);
  return 0;
}

extern "C" int raft_ss_plan_set_bp_variant(raft_ss_plan* p, int v) {
  if (!p) return ss_set_error(RAFT_SS_EINVAL, "null plan");
  int rc = check_uniform(p, "set_bp_variant");
  if (rc) return rc;
  rc = check_variant(p, v);
  if (rc) return rc;
  p->bp_variant = v;
  p->bp_variant_fixed = true;
  return 0;
}

extern "C" int raft_ss_plan_set_radon_slab_rows(raft_ss_plan* p, int rows) {
  if (!p) return ss_set_error(RAFT_SS_EINVAL, "null plan");
  if (rows < 0) return ss_set_error(RAFT_SS_EINVAL, "slab rows must be >= 0 (got %d)", rows);
  p->radon_rows_fixed = rows;
  return 0;
}

namespace {

struct EventPair {
  hipEvent_t a = nullptr, b = nullptr;
  ~EventPair() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  }
};

// best of `reps` timed runs of `run` (after one warm-up), in microseconds
template <class F>
int time_best(F run, int reps, hipStream_t s, double* best_us) {
  EventPair ev;
  HIP_TRY(hipEventCreate(&ev.a));
  HIP_TRY(hipEventCreate(&ev.b));
  int rc = run();
  if (rc) return rc;
  double best = 1e300;
  for (int r = 0; r < std::max(1, reps); ++r) {
    HIP_TRY(hipEventRecord(ev.a, s));
    if ((rc = run())) return rc;
    HIP_TRY(hipEventRecord(ev.b, s));
    HIP_TRY(hipEventSynchronize(ev.b));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, ev.a, ev.b));
    best = std::min(best, (double)ms * 1e3);
  }
  *best_us = best;
  return 0;
}

std::string device_arch() {
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return "unknown";
  return prop.gcnArchName;
}

struct TuneEntry {
  int variant = 0;
  double times[SS_BP_NVARIANTS];
};
std::mutex g_tune_mu;
std::map<std::string, TuneEntry> g_tune;

// ---- persisted choices
//
// The tuners' choices are kept per process in g_tune and in a JSON file, so that later
// processes on the same device architecture skip the timing: $PYRAFT_AMD_TUNE_CACHE (an
// empty value: no file), else $XDG_CACHE_HOME/pyraft_amd/tune.json, else
// $HOME/.cache/pyraft_amd/tune.json.  Keys name the architecture (gcnArchName) and the
// geometry (and the angle-range length of OSEM subsets).  One entry per line:
//   {"format": "<kTuneFormat>",
//    "entries": {
//     "<key>": {"variant": 3, "times_us": [t0, t1, t2, t3, t4, t5]},
//     ...
//    }}
// A file of another format is ignored and a line that does not parse is skipped (that
// geometry is timed again).  After each new choice the file is written whole to a
// temporary name and renamed over the old one.  Callers hold g_tune_mu.
const char* const kTuneFormat = "pyraft_amd tuning cache 3: backprojection variants x6, angle groups, radon shapes";
bool g_tune_loaded = false;

std::string tune_cache_path() {
  const char* env = getenv("PYRAFT_AMD_TUNE_CACHE");
  if (env) return env;  // "" disables the file
  const char* xdg = getenv("XDG_CACHE_HOME");
  if (xdg && *xdg) return std::string(xdg) + "/pyraft_amd/tune.json";
  const char* home = getenv("HOME");
  if (home && *home) return std::string(home) + "/.cache/pyraft_amd/tune.json";
  return "";
}

void tune_load_locked() {
  if (g_tune_loaded) return;
  g_tune_loaded = true;
  const std::string path = tune_cache_path();
  if (path.empty()) return;
  FILE* f = fopen(path.c_str(), "r");
  if (!f) return;
  char line[1024];
  bool format_ok = false;
  while (fgets(line, sizeof(line), f)) {
    if (!format_ok) {
      const char* q = strstr(line, "\"format\": \"");
      if (q) {
        q += strlen("\"format\": \"");
        const size_t len = strlen(kTuneFormat);
        format_ok = strncmp(q, kTuneFormat, len) == 0 && q[len] == '"';
        if (!format_ok) break;
      }
      continue;
    }
    const char* a = strchr(line, '"');
    const char* b = a ? strchr(a + 1, '"') : nullptr;
    if (!a || !b || b == a + 1) continue;
    TuneEntry te;
    double* t = te.times;
    if (sscanf(b + 1, ": {\"variant\": %d, \"times_us\": [%lf, %lf, %lf, %lf, %lf, %lf]}", &te.variant, &t[0], &t[1],
               &t[2], &t[3], &t[4], &t[5]) != 1 + SS_BP_NVARIANTS)
      continue;
    if (te.variant < 0 || te.variant >= SS_BP_NVARIANTS) continue;
    g_tune[std::string(a + 1, b)] = te;
  }
  fclose(f);
}

// mkdir -p of the directory part of `path`
void make_parent_dirs(const std::string& path) {
  for (size_t i = 1; i < path.size(); ++i)
    if (path[i] == '/') (void)mkdir(path.substr(0, i).c_str(), 0755);
}

void tune_save_locked() {
  const std::string path = tune_cache_path();
  if (path.empty()) return;
  make_parent_dirs(path);
  const std::string tmp = path + ".tmp." + std::to_string((long)getpid());
  FILE* f = fopen(tmp.c_str(), "w");
  if (!f) return;  // not writable: the choices stay per process
  fprintf(f, "{\"format\": \"%s\",\n \"entries\": {\n", kTuneFormat);
  size_t i = 0;
  for (const auto& kv : g_tune) {
    const double* t = kv.second.times;
    fprintf(f, "  \"%s\": {\"variant\": %d, \"times_us\": [%.3f, %.3f, %.3f, %.3f, %.3f, %.3f]}%s\n",
            kv.first.c_str(), kv.second.variant, t[0], t[1], t[2], t[3], t[4], t[5],
            ++i < g_tune.size() ? "," : "");
  }
  fprintf(f, " }}\n");
  const bool ok = fclose(f) == 0;
  if (!ok || rename(tmp.c_str(), path.c_str()) != 0) (void)remove(tmp.c_str());
}

// the cached choice for `key` (loading the file on first use)
bool tune_find(const std::string& key, TuneEntry* te) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  tune_load_locked();
  auto it = g_tune.find(key);
  if (it == g_tune.end()) return false;
  *te = it->second;
  return true;
}

void tune_store(const std::string& key, const TuneEntry& te) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  tune_load_locked();
  g_tune[key] = te;
  tune_save_locked();
}

std::string plan_key(const raft_ss_plan* p) {
  char buf[512];
  snprintf(buf, sizeof(buf), "%s|%d|%d|%d|%a|%a|%a|%a|%a|%a|%a|%a", device_arch().c_str(), p->g.n, p->g.nrays,
           p->g.nangles, p->img_corners[0], p->img_corners[1], p->img_corners[2], p->img_corners[3],
           p->sino_corners[0], p->sino_corners[1], p->sino_corners[2], p->sino_corners[3]);
  return buf;
}

}  // namespace

extern "C" int raft_ss_plan_tune(raft_ss_plan* p, int reps, int flags, double* times_us, void* stream_) {
  int rc = check_plan(p, 1);
  if (rc) return rc;
  if ((rc = check_uniform(p, "plan_tune"))) return rc;
  hipStream_t stream = as_stream(stream_);
  const std::string key = plan_key(p);
  TuneEntry cached;
  if (!(flags & RAFT_SS_TUNE_FORCE) && tune_find(key, &cached) && p->bp_angles_v[cached.variant] > 0) {
    p->bp_variant = cached.variant;
    if (times_us) std::copy(cached.times, cached.times + SS_BP_NVARIANTS, times_us);
    return p->bp_variant;
  }
  const int n = p->g.n, R = p->g.nrays, A = p->g.nangles;
  size_t part_bytes = 0;  // angle-group partial images of the variants that divide their launch
  for (int v = 0; v < SS_BP_NVARIANTS; ++v)
    if (p->bp_angles_v[v] > 0) part_bytes = std::max(part_bytes, ss_bp_scratch(p->dev(v), 0, A, 1));
  DevBuf ws;
  HIP_TRY(ws.alloc(sizeof(float) * ((size_t)R * A + (size_t)n * n) + part_bytes, stream));
  float* sinoT = ws.as<float>();
  float* img = sinoT + (size_t)R * A;
  float* part = img + (size_t)n * n;
  HIP_TRY(ss_launch_fill(sinoT, 1.0f, (size_t)R * A, stream));
  TuneEntry te;
  int best = -1;
  for (int v = 0; v < SS_BP_NVARIANTS; ++v) {
    te.times[v] = -1.0;
    if (p->bp_angles_v[v] <= 0) continue;
    const SSDevPlan dp = p->dev(v);
    SSEmUpdateArgs none{nullptr, nullptr};
    auto run = [&]() -> int {
      HIP_TRY(ss_launch_backprojection(dp, sinoT, 0, A, 1, SS_BP_PLAIN, img, none, part, stream));
      return 0;
    };
    if ((rc = time_best(run, reps, stream, &te.times[v]))) return rc;
    if (best < 0 || te.times[v] < te.times[best]) best = v;
  }
  if (best < 0) return ss_set_error(RAFT_SS_EINTERNAL, "no runnable backprojection variant");
  te.variant = best;
  p->bp_variant = best;
  if (times_us) std::copy(te.times, te.times + SS_BP_NVARIANTS, times_us);
  tune_store(key, te);
  return best;
}

extern "C" int raft_ss_plan_tune_transpose(raft_ss_plan* p, int nslices, int reps, int flags, double* times_us,
                                           void* stream_) {
  int rc = check_plan(p, std::max(nslices, 1));
  if (rc) return rc;
  if ((rc = check_uniform(p, "plan_tune_transpose"))) return rc;
  if (nslices < 1) return ss_set_error(RAFT_SS_EINVAL, "nslices must be >= 1");
  if (times_us)
    for (int v = 0; v < SS_BP_NVARIANTS; ++v) times_us[v] = -1.0;
  if (p->rt_variant < 0) return -1;  // line form: nothing to tune
  hipStream_t stream = as_stream(stream_);
  const std::string key = plan_key(p) + "|transpose|s" + std::to_string(nslices);
  TuneEntry cached;
  if (!(flags & RAFT_SS_TUNE_FORCE) && tune_find(key, &cached) && p->rt_angles_v[cached.variant] > 0) {
    p->rt_variant = cached.variant;
    if (times_us) std::copy(cached.times, cached.times + SS_BP_NVARIANTS, times_us);
    return p->rt_variant;
  }
  const int n = p->g.n, R = p->g.nrays, A = p->g.nangles;
  DevBuf ws;
  HIP_TRY(ws.alloc(sizeof(float) * (size_t)nslices * (2 * (size_t)R * A + (size_t)n * n), stream));
  float* sinoT = ws.as<float>();
  float* img = sinoT + (size_t)nslices * R * A;
  float* scratch = img + (size_t)nslices * n * n;  // the pre-scaled sinogram
  HIP_TRY(ss_launch_fill(sinoT, 1.0f, (size_t)nslices * R * A, stream));
  TuneEntry te;
  int best = -1;
  const int keep = p->rt_variant;
  for (int v = 0; v < SS_BP_NVARIANTS; ++v) {
    te.times[v] = -1.0;
    if (p->rt_angles_v[v] <= 0) continue;
    p->rt_variant = v;
    const SSDevPlan dp = p->dev();
    auto run = [&]() -> int {
      HIP_TRY(ss_launch_radon_transpose(dp, sinoT, 0, A, nslices, img, scratch, stream));
      return 0;
    };
    rc = time_best(run, reps, stream, &te.times[v]);
    p->rt_variant = keep;
    if (rc) return rc;
    if (best < 0 || te.times[v] < te.times[best]) best = v;
  }
  if (best < 0) return ss_set_error(RAFT_SS_EINTERNAL, "no runnable transpose variant");
  te.variant = best;
  p->rt_variant = best;
  if (times_us) std::copy(te.times, te.times + SS_BP_NVARIANTS, times_us);
  tune_store(key, te);
  return best;
}

int ss_plan_range_variant(raft_ss_plan* p, int j0, int j1, hipStream_t stream, int* variant) {
  *variant = p->bp_variant;
  if (p->bp_variant_fixed || j1 <= j0) return 0;
  const std::string key = plan_key(p) + "|range" + std::to_string(j1 - j0);
  TuneEntry cached;
  if (tune_find(key, &cached) && p->bp_angles_v[cached.variant] > 0) {
    *variant = cached.variant;
    return 0;
  }
  // no timing inside a stream capture: the plan's variant until an eager call tunes
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  HIP_TRY(hipStreamIsCapturing(stream, &cs));
  if (cs != hipStreamCaptureStatusNone) return 0;
  const int n = p->g.n, R = p->g.nrays, A = p->g.nangles;
  size_t part_bytes = 0;
  for (int v = 0; v < SS_BP_NVARIANTS; ++v)
    if (p->bp_angles_v[v] > 0) part_bytes = std::max(part_bytes, ss_bp_scratch(p->dev(v), j0, j1, 1));
  DevBuf ws;
  HIP_TRY(ws.alloc(sizeof(float) * ((size_t)R * A + (size_t)n * n) + part_bytes, stream));
  float* sinoT = ws.as<float>();
  float* img = sinoT + (size_t)R * A;
  float* part = img + (size_t)n * n;
  HIP_TRY(ss_launch_fill(sinoT, 1.0f, (size_t)R * A, stream));
  TuneEntry te;
  int best = -1, rc = 0;
  for (int v = 0; v < SS_BP_NVARIANTS; ++v) {
    te.times[v] = -1.0;
    if (p->bp_angles_v[v] <= 0) continue;
    const SSDevPlan dp = p->dev(v);
    SSEmUpdateArgs none{nullptr, nullptr};
    auto run = [&]() -> int {
      HIP_TRY(ss_launch_backprojection(dp, sinoT, j0, j1, 1, SS_BP_PLAIN, img, none, part, stream));
      return 0;
    };
    if ((rc = time_best(run, 3, stream, &te.times[v]))) return rc;
    if (best < 0 || te.times[v] < te.times[best]) best = v;
  }
  if (best < 0) return ss_set_error(RAFT_SS_EINTERNAL, "no runnable backprojection variant");
  te.variant = best;
  *variant = best;
  tune_store(key, te);
  return 0;
}

int ss_plan_radon_rows(const raft_ss_plan* p, int j0, int j1, int nslices, hipStream_t stream, bool may_time,
                       SSDevPlan* out) {
  out->radon_slab_rows = 0;
  out->radon_shape = 0;
  const int n = p->g.n;
  if (p->radon_rows_fixed > 0) {  // raft_ss_plan_set_radon_slab_rows
    out->radon_slab_rows = p->radon_rows_fixed;
    return 0;
  }
  if (j1 <= j0) return 0;
  SSDevPlan dp = p->dev();
  int r0 = n;
  const bool divides = ss_radon_slabs(dp, j0, j1, nslices, &r0) > 1;
  const long long pk = ((long long)(j1 - j0) << 20) | (long long)std::min(nslices, (1 << 20) - 1);
  {
    std::lock_guard<std::mutex> lock(p->radon_mutex);
    auto it = p->radon_rows.find(pk);
    if (it != p->radon_rows.end()) {
      out->radon_slab_rows = it->second & 0xffffff;
      out->radon_shape = it->second >> 24;
      return 0;
    }
  }
  if (!may_time) return 0;  // the rule, shape 0 (no lookups that allocate)
  // candidates (slab rows, shape); rows 0 is the default rule, which does not divide here
  std::vector<std::pair<int, int>> cand;
  if (divides) {
    cand = {{r0, 0}, {r0, 1}};
    if (2 * r0 < n) cand.push_back({2 * r0, 0});
    if (r0 / 2 >= RADON_SLAB_ALIGN && (r0 / 2) % RADON_SLAB_ALIGN == 0) cand.push_back({r0 / 2, 0});
    cand.push_back({n, 0});
    cand.push_back({n, 1});
  } else {
    cand = {{0, 0}, {0, 1}};
  }
  const std::string key = plan_key(p) + "|radon|range" + std::to_string(j1 - j0) + "|s" + std::to_string(nslices);
  // mesh plans are timed per plan but kept out of the shared cache: its key does not
  // name their t and theta tables
  const bool shared = !p->mesh;
  TuneEntry cached;
  int chosen = -1;
  if (shared && tune_find(key, &cached) && cached.variant >= 0 && cached.variant < (int)cand.size())
    chosen = cached.variant;
  if (chosen < 0) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    HIP_TRY(hipStreamIsCapturing(stream, &cs));
    if (cs != hipStreamCaptureStatusNone) return 0;  // the rule, untimed
    const size_t layout = (size_t)nslices * n * p->g.pitch;
    size_t slab_bytes = 0;
    for (const auto& c : cand) {
      dp.radon_slab_rows = c.first;
      slab_bytes = std::max(slab_bytes, ss_radon_scratch(dp, j0, j1, nslices));
    }
    const size_t img = (size_t)nslices * n * n, sino = (size_t)nslices * p->g.nrays * p->g.nangles;
    DevBuf ws;
    HIP_TRY(ws.alloc(2 * layout * sizeof(float2) + (img + sino) * sizeof(float) + slab_bytes, stream));
    float2* gA = ws.as<float2>();
    float2* gB = gA + layout;
    float* x = reinterpret_cast<float*>(gB + layout);
    float* sout = x + img;
    float* slabs = sout + sino;
    HIP_TRY(ss_launch_fill(x, 1.0f, img, stream));
    HIP_TRY(ss_launch_gather_layouts(dp, x, gA, gB, nslices, stream));
    TuneEntry te;
    for (int v = 0; v < SS_BP_NVARIANTS; ++v) te.times[v] = -1.0;
    int best = 0, rc = 0;
    for (int c = 0; c < (int)cand.size() && c < SS_BP_NVARIANTS; ++c) {
      dp.radon_slab_rows = cand[c].first;
      dp.radon_shape = cand[c].second;
      SSEmRatioArgs none{nullptr, nullptr};
      auto run = [&]() -> int {
        HIP_TRY(ss_launch_radon(dp, gA, gB, j0, j1, nslices, SS_OUT_SINOT, sout, none, slabs, stream));
        return 0;
      };
      if ((rc = time_best(run, 3, stream, &te.times[c]))) return rc;
      if (te.times[c] < te.times[best]) best = c;
    }
    te.variant = best;
    if (shared) tune_store(key, te);
    chosen = best;
  }
  std::lock_guard<std::mutex> lock(p->radon_mutex);
  p->radon_rows[pk] = cand[chosen].first | (cand[chosen].second << 24);
  out->radon_slab_rows = cand[chosen].first;
  out->radon_shape = cand[chosen].second;
  return 0;
}

// ------------------------------------------------------------------ batches

namespace {

struct BatchLaunch {
  std::vector<SSBatchEntry> entries;
  SSBatchEntry* d_entries = nullptr;
  int total = 0, grid = 0;
  bool slabs = false;          // Radon / finish: the launch works in row slabs (partial sums)
  size_t part_floats = 0;      // Radon / finish: partial sums of every entry
  std::vector<int> entry_of;   // slice -> entry index (-1: not in this launch)
  BatchLaunch() = default;
  BatchLaunch(const BatchLaunch&) = delete;
  BatchLaunch& operator=(const BatchLaunch&) = delete;
  ~BatchLaunch() {
    if (d_entries) (void)hipFree(d_entries);
  }
  SSBatchLaunch dev(const SSBatchSlice* d_slices) const {
    SSBatchLaunch L;
    L.slices = d_slices;
    L.entries = d_entries;
    L.nentries = (int)entries.size();
    L.total = total;
    L.grid = grid;
    return L;
  }
};

// grow-only device buffer owned by the batch
struct Buf {
  void* p = nullptr;
  size_t bytes = 0;
  ~Buf() {
    if (p) (void)hipFree(p);
  }
  // returns 1 if the address changed
  int ensure(size_t want, hipError_t* err) {
    *err = hipSuccess;
    if (want <= bytes && p) return 0;
    if (p) {
      *err = hipFree(p);
      p = nullptr;
      bytes = 0;
      if (*err != hipSuccess) return 1;
    }
    *err = hipMalloc(&p, want ? want : 256);
    if (*err == hipSuccess) bytes = want;
    return 1;
  }
  template <class T>
  T* as(size_t byte_off = 0) const {
    return reinterpret_cast<T*>(reinterpret_cast<char*>(p) + byte_off);
  }
};

struct GraphEntry {
  std::string key;
  hipGraphExec_t exec = nullptr;
};

int64_t align_up(int64_t v, int64_t a) { return (v + a - 1) / a * a; }

}  // namespace

struct raft_ss_batch {
  int device = 0;
  int nslices = 0;
  std::vector<std::shared_ptr<raft_ss_plan>> plans;  // per slice (shared by equal geometries)
  std::vector<int64_t> img_off, sino_off, lay_off;
  int64_t img_total = 0, sino_total = 0, lay_total = 0;
  std::vector<int> variant;  // backprojection variant per slice
  int gather_tile = SS_GATHER_TILE;  // gather-layout tile side (ss_batch_gather_tile)
  Buf slices;                // SSBatchSlice[nslices]
  bool slices_dirty = true;
  std::map<std::string, std::unique_ptr<BatchLaunch>> launches;
  Buf lay, t0, t1, weights, partials, ll, ranges, slab;
  Buf em_ws;  // a one-slice batch's OSEM workspace (the slice's own kernels)
  int generation = 0;        // bumped whenever a buffer moves or the variants change
  int weights_S = 0;         // subsets the cached BP_b(1) images were computed for
  int weights_generation = -1;
  std::vector<GraphEntry> graphs;
  hipStream_t capture = nullptr;
  std::mutex mu;

  ~raft_ss_batch() {
    for (auto& g : graphs)
      if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (capture) (void)hipStreamDestroy(capture);
  }
  const raft_ss_plan& plan(int i) const { return *plans[i]; }
};

namespace {

void drop_graphs(raft_ss_batch* b) {
  for (auto& g : b->graphs)
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
  b->graphs.clear();
}

// a device buffer of `bytes`; a move invalidates graphs and cached weights
int ensure_buf(raft_ss_batch* b, Buf& buf, size_t bytes) {
  hipError_t e;
  if (buf.ensure(bytes, &e)) {
    ++b->generation;
    drop_graphs(b);
  }
  HIP_TRY(e);
  return 0;
}

int sync_slices(raft_ss_batch* b) {
  if (!b->slices_dirty) return 0;
  std::vector<SSBatchSlice> h(b->nslices);
  for (int i = 0; i < b->nslices; ++i) {
    h[i].p = b->plan(i).dev(b->variant[i]);
    h[i].img_off = b->img_off[i];
    h[i].sino_off = b->sino_off[i];
    h[i].lay_off = b->lay_off[i];
  }
  int rc = ensure_buf(b, b->slices, sizeof(SSBatchSlice) * h.size());
  if (rc) return rc;
  HIP_TRY(hipMemcpy(b->slices.p, h.data(), sizeof(SSBatchSlice) * h.size(), hipMemcpyHostToDevice));
  b->slices_dirty = false;
  return 0;
}

// subset b of S of slice i: angles [b*A/S, (b+1)*A/S) (raft_ss_em's partition)
void subset_range(int A, int S, int bsub, int* j0, int* j1) {
  *j0 = (int)(((long long)bsub * A) / S);
  *j1 = (int)(((long long)(bsub + 1) * A) / S);
}

// Items per entry and cost per item of each launch kind; entries sorted by
// decreasing cost per item so the cheapest items run last.  A Radon launch (and its
// finishing pass) of fewer than SS_RADON_SLAB_BLOCKS blocks divides every entry's
// sample rows into slabs; the entries' partial sums are laid out in slice order.
int build_launch(raft_ss_batch* b, BatchLaunch& L, int kind, int S, int bsub, int mode, int v) {
  struct Cand {
    SSBatchEntry e;
    int items;
    double cost;
  };
  std::vector<Cand> c;
  long long blocks = 0;
  for (int i = 0; i < b->nslices; ++i) {
    const raft_ss_plan& p = b->plan(i);
    int j0 = 0, j1 = p.g.nangles;
    if (S > 0) subset_range(p.g.nangles, S, bsub, &j0, &j1);
    blocks += ss_radon_blocks(p.nwaves, j0, j1);
  }
  const bool radon = kind == SS_KIND_RADON || kind == SS_KIND_RADON_FINISH;
  L.slabs = radon && blocks < SS_RADON_SLAB_BLOCKS;
  L.part_floats = 0;
  for (int i = 0; i < b->nslices; ++i) {
    const raft_ss_plan& p = b->plan(i);
    const int n = p.g.n, R = p.g.nrays, A = p.g.nangles;
    int j0 = 0, j1 = A;
    if (S > 0) subset_range(A, S, bsub, &j0, &j1);
    Cand k;
    k.e.slice = i;
    k.e.j0 = j0;
    k.e.j1 = j1;
    k.e.item0 = 0;
    k.e.nslab = 1;
    k.e.slab_rows = n;
    k.e.part_off = 0;
    k.e.which = SS_GATHER_BOTH;
    if (kind == SS_KIND_GATHER && S > 0) {  // a subset's gather: the layouts its angles read
      k.e.which = 0;
      for (int j = j0; j < j1; ++j) k.e.which |= p.h_angles[j].branch_a ? SS_GATHER_A : SS_GATHER_B;
    }
    if (L.slabs && j1 > j0) {
      k.e.nslab = ss_radon_slab_plan(n, blocks, &k.e.slab_rows);
      k.e.part_off = (int64_t)L.part_floats;
      L.part_floats += (size_t)k.e.nslab * (size_t)(j1 - j0) * R;
    }
    switch (kind) {
      case SS_KIND_GATHER: {
        const int t = (n + 2 + b->gather_tile - 1) / b->gather_tile;
        k.items = t * t;
        k.cost = 1.0;
        break;
      }
      case SS_KIND_TRANSPOSE:
        k.items = ((R + 31) / 32) * ((A + 31) / 32);
        k.cost = 1.0;
        break;
      case SS_KIND_RADON:
        k.items = (int)ss_radon_blocks(p.nwaves, j0, j1) * k.e.nslab;
        k.cost = (double)k.e.slab_rows * 8;
        break;
      case SS_KIND_RADON_FINISH:  // only in slab form: 256 rays per item
        k.items = L.slabs ? (int)(((long long)(j1 - j0) * R + 255) / 256) : 0;
        k.cost = (double)k.e.nslab;
        break;
      case SS_KIND_FIXUP:
        k.items = p.nirregular;
        k.cost = (double)n;
        break;
      default:  // SS_KIND_BP, only the slices of variant v
        if (b->variant[i] != v) continue;
        k.items = j1 > j0 ? ss_bp_tiles(n, v) : 0;
        k.cost = (double)(j1 - j0) * kSSBpRows[v];
        break;
    }
    if (k.items > 0) c.push_back(k);
  }
  std::stable_sort(c.begin(), c.end(), [](const Cand& x, const Cand& y) { return x.cost > y.cost; });
  L.entries.clear();
  L.entry_of.assign(b->nslices, -1);
  long long total = 0;
  for (const Cand& k : c) {
    SSBatchEntry e = k.e;
    e.item0 = (int)total;
    L.entry_of[e.slice] = (int)L.entries.size();
    L.entries.push_back(e);
    total += k.items;
  }
  if (total > 0x7fffffffLL) return ss_set_error(RAFT_SS_EINVAL, "batch too large (%lld work items)", total);
  L.total = (int)total;
  L.grid = ss_batch_grid(kind, v, kind == SS_KIND_RADON && L.slabs ? SS_OUT_PARTIAL : mode, L.total);
  if (!L.entries.empty()) {
    HIP_TRY(hipMalloc(&L.d_entries, sizeof(SSBatchEntry) * L.entries.size()));
    HIP_TRY(hipMemcpy(L.d_entries, L.entries.data(), sizeof(SSBatchEntry) * L.entries.size(),
                      hipMemcpyHostToDevice));
  }
  return 0;
}

// cached launch table (built on first use; never during graph capture)
int get_launch(raft_ss_batch* b, BatchLaunch** out, int kind, int S, int bsub, int mode, int v) {
  char key[96];
  snprintf(key, sizeof(key), "%d|%d|%d|%d|%d", kind, S, bsub, mode, kind == SS_KIND_BP ? v : -1);
  auto it = b->launches.find(key);
  if (it == b->launches.end()) {
    std::unique_ptr<BatchLaunch> L(new BatchLaunch());
    int rc = build_launch(b, *L, kind, S, bsub, mode, v);
    if (rc) return rc;
    it = b->launches.emplace(key, std::move(L)).first;
  }
  *out = it->second.get();
  return 0;
}

// variants present in the batch
std::vector<int> variants_used(const raft_ss_batch* b) {
  std::vector<int> v(b->variant);
  std::sort(v.begin(), v.end());
  v.erase(std::unique(v.begin(), v.end()), v.end());
  return v;
}

void variants_changed(raft_ss_batch* b) {
  b->slices_dirty = true;
  for (auto it = b->launches.begin(); it != b->launches.end();) {
    if (it->first.rfind(std::to_string(SS_KIND_BP) + "|", 0) == 0)
      it = b->launches.erase(it);
    else
      ++it;
  }
  ++b->generation;
  drop_graphs(b);
}

int check_batch(const raft_ss_batch* b) {
  if (!b) return ss_set_error(RAFT_SS_EINVAL, "null batch");
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev != b->device)
    return ss_set_error(RAFT_SS_EINVAL, "batch was built on device %d but device %d is current", b->device, dev);
  return 0;
}

// Run `body` on `user` directly, or (graph) capture it once on the batch's own
// stream under `key` and launch the instantiated graph on `user`.
template <class F>
int run_body(raft_ss_batch* b, bool graph, const std::string& key, hipStream_t user, F body) {
  if (!graph) return body(user);
  const std::string full = key + "|g" + std::to_string(b->generation);
  hipGraphExec_t exec = nullptr;
  for (auto& g : b->graphs)
    if (g.key == full) exec = g.exec;
  if (!exec) {
    if (!b->capture) HIP_TRY(hipStreamCreateWithFlags(&b->capture, hipStreamNonBlocking));
    HIP_TRY(hipStreamBeginCapture(b->capture, hipStreamCaptureModeThreadLocal));
    const int rc = body(b->capture);
    hipGraph_t graph_ = nullptr;
    const hipError_t e = hipStreamEndCapture(b->capture, &graph_);
    if (rc) {
      if (graph_) (void)hipGraphDestroy(graph_);
      return rc;
    }
    HIP_TRY(e);
    const hipError_t ei = hipGraphInstantiate(&exec, graph_, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph_);
    HIP_TRY(ei);
    if (b->graphs.size() >= 8) {
      (void)hipGraphExecDestroy(b->graphs.front().exec);
      b->graphs.erase(b->graphs.begin());
    }
    b->graphs.push_back(GraphEntry{full, exec});
  }
  HIP_TRY(hipGraphLaunch(exec, user));
  return 0;
}

}  // namespace

extern "C" int raft_ss_batch_create(raft_ss_batch** out, int nslices, const int* n, const int* nrays,
                                    const int* nangles, const double* img_corners, const double* sino_corners,
                                    void* stream) {
  if (!out || !n || !nrays || !nangles) return ss_set_error(RAFT_SS_EINVAL, "null argument");
  *out = nullptr;
  if (nslices < 1) return ss_set_error(RAFT_SS_EINVAL, "nslices must be >= 1 (got %d)", nslices);
  std::unique_ptr<raft_ss_batch> b(new raft_ss_batch());
  HIP_TRY(hipGetDevice(&b->device));
  b->nslices = nslices;
  const double dic[4] = {-1.0, 1.0, 1.0, -1.0};
  const double dsc[4] = {0.0, 1.0, 3.141592653589793, -1.0};
  std::map<std::string, std::shared_ptr<raft_ss_plan>> by_geom;
  b->plans.resize(nslices);
  b->variant.resize(nslices);
  b->img_off.resize(nslices);
  b->sino_off.resize(nslices);
  b->lay_off.resize(nslices);
  for (int i = 0; i < nslices; ++i) {
    const double* ic = img_corners ? img_corners + 4 * i : dic;
    const double* sc = sino_corners ? sino_corners + 4 * i : dsc;
    char key[256];
    snprintf(key, sizeof(key), "%d|%d|%d|%a|%a|%a|%a|%a|%a|%a|%a", n[i], nrays[i], nangles[i], ic[0], ic[1], ic[2],
             ic[3], sc[0], sc[1], sc[2], sc[3]);
    auto it = by_geom.find(key);
    if (it == by_geom.end()) {
      raft_ss_plan* p = nullptr;
      int rc = raft_ss_plan_create(&p, n[i], nrays[i], nangles[i], ic, sc, stream);
      if (rc) return ss_set_error(rc, "slice %d: %s", i, raft_ss_last_error());
      it = by_geom.emplace(key, std::shared_ptr<raft_ss_plan>(p, raft_ss_plan_destroy)).first;
    }
    b->plans[i] = it->second;
    b->variant[i] = it->second->bp_variant;
    // 256-byte aligned slices in the packed buffers
    b->img_off[i] = b->img_total;
    b->img_total = align_up(b->img_total + (int64_t)n[i] * n[i], 64);
    b->sino_off[i] = b->sino_total;
    b->sino_total = align_up(b->sino_total + (int64_t)nrays[i] * nangles[i], 64);
    b->lay_off[i] = b->lay_total;
    b->lay_total = align_up(b->lay_total + (int64_t)n[i] * it->second->g.pitch, 32);
  }
  b->gather_tile = ss_batch_gather_tile(n, nslices);
  *out = b.release();
  return 0;
}

extern "C" int raft_ss_batch_destroy(raft_ss_batch* b) {
  delete b;
  return 0;
}

extern "C" int raft_ss_batch_layout(const raft_ss_batch* b, long long* img_off, long long* sino_off,
                                    long long* totals) {
  if (!b) return ss_set_error(RAFT_SS_EINVAL, "null batch");
  for (int i = 0; i < b->nslices; ++i) {
    if (img_off) img_off[i] = b->img_off[i];
    if (sino_off) sino_off[i] = b->sino_off[i];
  }
  if (totals) {
    totals[0] = b->img_total;
    totals[1] = b->sino_total;
  }
  return 0;
}

extern "C" int raft_ss_batch_info(const raft_ss_batch* b, int slice, int* info, int ninfo) {
  if (!b || !info) return ss_set_error(RAFT_SS_EINVAL, "null argument");
  if (slice < 0 || slice >= b->nslices)
    return ss_set_error(RAFT_SS_EINVAL, "slice %d outside [0, %d)", slice, b->nslices);
  const raft_ss_plan& p = b->plan(slice);
  const int v = b->variant[slice];
  const int vals[] = {p.g.n, p.g.nrays, p.g.nangles, p.nirregular, v, p.bp_angles_v[v]};
  for (int i = 0; i < ninfo && i < (int)(sizeof(vals) / sizeof(vals[0])); ++i) info[i] = vals[i];
  return 0;
}

extern "C" int raft_ss_batch_set_bp_variant(raft_ss_batch* b, int slice, int v) {
  if (!b) return ss_set_error(RAFT_SS_EINVAL, "null batch");
  std::lock_guard<std::mutex> lock(b->mu);
  const int i0 = slice < 0 ? 0 : slice, i1 = slice < 0 ? b->nslices : slice + 1;
  if (slice >= b->nslices) return ss_set_error(RAFT_SS_EINVAL, "slice %d outside [0, %d)", slice, b->nslices);
  for (int i = i0; i < i1; ++i) {
    int rc = check_variant(&b->plan(i), v);
    if (rc) return ss_set_error(rc, "slice %d: %s", i, raft_ss_last_error());
  }
  for (int i = i0; i < i1; ++i) b->variant[i] = v;
  variants_changed(b);
  return 0;
}

extern "C" int raft_ss_batch_radon(raft_ss_batch* b, const float* d_images, float* d_sinos, int angle_major,
                                   int flags, void* stream_) {
  int rc = check_batch(b);
  if (rc) return rc;
  if (!d_images || !d_sinos) return ss_set_error(RAFT_SS_EINVAL, "null device pointer");
  std::lock_guard<std::mutex> lock(b->mu);
  hipStream_t stream = as_stream(stream_);
  const int mode = angle_major ? SS_OUT_SINOT : SS_OUT_SINO;
  BatchLaunch *Lg, *Lr, *Lz, *Lf;
  if ((rc = sync_slices(b)) || (rc = get_launch(b, &Lg, SS_KIND_GATHER, 0, 0, 0, 0)) ||
      (rc = get_launch(b, &Lr, SS_KIND_RADON, 0, 0, mode, 0)) ||
      (rc = get_launch(b, &Lz, SS_KIND_RADON_FINISH, 0, 0, mode, 0)) ||
      (rc = get_launch(b, &Lf, SS_KIND_FIXUP, 0, 0, mode, 0)) ||
      (rc = ensure_buf(b, b->lay, sizeof(float2) * 2 * (size_t)b->lay_total)) ||
      (rc = ensure_buf(b, b->slab, sizeof(float) * std::max<size_t>(Lr->part_floats, 1))))
    return rc;
  float2* gA = b->lay.as<float2>();
  float2* gB = gA + b->lay_total;
  float* part = Lr->slabs ? b->slab.as<float>() : nullptr;
  const SSBatchSlice* ds = b->slices.as<SSBatchSlice>();
  char key[128];
  snprintf(key, sizeof(key), "radon|%p|%p|%d", (const void*)d_images, (void*)d_sinos, mode);
  return run_body(b, flags & RAFT_SS_BATCH_GRAPH, key, stream, [&](hipStream_t s) -> int {
    HIP_TRY(ss_launch_batch_gather(Lg->dev(ds), b->gather_tile, d_images, gA, gB, s));
    HIP_TRY(ss_launch_batch_radon(Lr->dev(ds), gA, gB, part, mode, d_sinos, nullptr, nullptr, s));
    if (part) HIP_TRY(ss_launch_batch_radon_finish(Lz->dev(ds), part, mode, d_sinos, nullptr, nullptr, s));
    HIP_TRY(ss_launch_batch_radon_fixup(Lf->dev(ds), gA, gB, mode, d_sinos, nullptr, nullptr, s));
    return 0;
  });
}

extern "C" int raft_ss_batch_backprojection(raft_ss_batch* b, const float* d_sinos, float* d_images,
                                            int angle_major, int flags, void* stream_) {
  int rc = check_batch(b);
  if (rc) return rc;
  if (!d_images || !d_sinos) return ss_set_error(RAFT_SS_EINVAL, "null device pointer");
  std::lock_guard<std::mutex> lock(b->mu);
  hipStream_t stream = as_stream(stream_);
  BatchLaunch* Lt = nullptr;
  if ((rc = sync_slices(b)) || (rc = get_launch(b, &Lt, SS_KIND_TRANSPOSE, 0, 0, 0, 0))) return rc;
  const std::vector<int> vs = variants_used(b);
  std::vector<BatchLaunch*> Lb(vs.size());
  for (size_t q = 0; q < vs.size(); ++q)
    if ((rc = get_launch(b, &Lb[q], SS_KIND_BP, 0, 0, SS_BP_PLAIN, vs[q]))) return rc;
  if (!angle_major && (rc = ensure_buf(b, b->t0, sizeof(float) * (size_t)b->sino_total))) return rc;
  const SSBatchSlice* ds = b->slices.as<SSBatchSlice>();
  float* sT = angle_major ? const_cast<float*>(d_sinos) : b->t0.as<float>();
  char key[128];
  snprintf(key, sizeof(key), "bp|%p|%p|%d", (const void*)d_sinos, (void*)d_images, angle_major);
  return run_body(b, flags & RAFT_SS_BATCH_GRAPH, key, stream, [&](hipStream_t s) -> int {
    if (!angle_major) HIP_TRY(ss_launch_batch_transpose(Lt->dev(ds), d_sinos, sT, 1, s));
    for (size_t q = 0; q < vs.size(); ++q)
      HIP_TRY(ss_launch_batch_backprojection(Lb[q]->dev(ds), vs[q], sT, SS_BP_PLAIN, d_images, nullptr, nullptr, s));
    return 0;
  });
}

extern "C" int raft_ss_batch_em(raft_ss_batch* b, const float* d_sinos, float* d_x, int niter, int nsubsets,
                                double* ll_host, int flags, void* stream_) {
  int rc = check_batch(b);
  if (rc) return rc;
  if (!d_sinos || !d_x) return ss_set_error(RAFT_SS_EINVAL, "null device pointer");
  if (niter < 0) return ss_set_error(RAFT_SS_EINVAL, "niter must be >= 0");
  int amin = 1 << 30;
  for (int i = 0; i < b->nslices; ++i) amin = std::min(amin, b->plan(i).g.nangles);
  if (nsubsets < 1 || nsubsets > amin)
    return ss_set_error(RAFT_SS_EINVAL, "nsubsets must be in [1, %d] (the smallest angle count)", amin);
  std::lock_guard<std::mutex> lock(b->mu);
  hipStream_t stream = as_stream(stream_);
  const int S = nsubsets, ns = b->nslices;
  if (ns == 1) {
    // One slice leaves nothing to pack into shared launches: the slice's own OSEM
    // kernels (raft_ss_em's, sized and tuned for its subsets) run instead, captured
    // into the batch's graph like the batched kernels, with the LL in the batch buffer.
    const raft_ss_plan* p0 = &b->plan(0);
    SSEmLaunch L;
    if ((rc = ss_em_prepare(p0, S, 1, niter, stream, &L)) ||
        (rc = ensure_buf(b, b->em_ws, std::max<size_t>(L.bytes, 256))) ||
        (rc = ensure_buf(b, b->ll, sizeof(double) * (size_t)std::max(niter, 1))))
      return rc;
    char key1[160];
    snprintf(key1, sizeof(key1), "em1|%p|%p|%d|%d", (const void*)d_sinos, (void*)d_x, niter, S);
    double* lld = b->ll.as<double>();
    void* ws = b->em_ws.p;
    rc = run_body(b, flags & RAFT_SS_BATCH_GRAPH, key1, stream, [&](hipStream_t st) -> int {
      return ss_em_launch(p0, L, d_sinos, d_x, niter, 1, ws, lld, st);
    });
    if (rc) return rc;
    if (ll_host && niter > 0) {
      HIP_TRY(hipMemcpyAsync(ll_host, lld, sizeof(double) * (size_t)niter, hipMemcpyDeviceToHost, stream));
      HIP_TRY(hipStreamSynchronize(stream));
    }
    return 0;
  }
  const std::vector<int> vs = variants_used(b);

  // ---- everything the captured part needs, before any capture
  BatchLaunch* Lt;
  if ((rc = sync_slices(b)) || (rc = get_launch(b, &Lt, SS_KIND_TRANSPOSE, 0, 0, 0, 0))) return rc;
  std::vector<BatchLaunch*> Lg(S), Lr(S), Lz(S), Lm(S), Lf(S), Lbw((size_t)S * vs.size()),
      Lbu((size_t)S * vs.size());
  std::vector<long long> poff(S + 1, 0);
  size_t part_floats = 1;
  for (int s = 0; s < S; ++s) {
    if ((rc = get_launch(b, &Lg[s], SS_KIND_GATHER, S, s, 0, 0)) ||
        (rc = get_launch(b, &Lr[s], SS_KIND_RADON, S, s, SS_OUT_EM_RATIO, 0)) ||
        (rc = get_launch(b, &Lz[s], SS_KIND_RADON_FINISH, S, s, SS_OUT_EM_RATIO, 0)) ||
        (rc = get_launch(b, &Lf[s], SS_KIND_FIXUP, S, s, SS_OUT_EM_RATIO, 0)))
      return rc;
    for (size_t q = 0; q < vs.size(); ++q)
      if ((rc = get_launch(b, &Lbw[s * vs.size() + q], SS_KIND_BP, S, s, SS_BP_PLAIN, vs[q])) ||
          (rc = get_launch(b, &Lbu[s * vs.size() + q], SS_KIND_BP, S, s, SS_BP_EM_UPDATE, vs[q])))
        return rc;
    // the launch whose items write the LL partials: the finishing pass in slab form
    Lm[s] = Lr[s]->slabs ? Lz[s] : Lr[s];
    if (Lr[s]->slabs) part_floats = std::max(part_floats, Lr[s]->part_floats);
    poff[s + 1] = poff[s] + Lm[s]->total + Lf[s]->total;
  }
  // per-slice (start, count) partial ranges: radon (or finish) then fixup items of every subset
  std::vector<int2> ranges((size_t)ns * 2 * S);
  for (int i = 0; i < ns; ++i)
    for (int s = 0; s < S; ++s) {
      int2 r0{0, 0}, r1{0, 0};
      const BatchLaunch& M = *Lm[s];
      const int er = M.entry_of[i], ef = Lf[s]->entry_of[i];
      if (er >= 0) {
        const int next = er + 1 < (int)M.entries.size() ? M.entries[er + 1].item0 : M.total;
        r0 = int2{(int)(poff[s] + M.entries[er].item0), next - M.entries[er].item0};
      }
      if (ef >= 0) {
        const int next = ef + 1 < (int)Lf[s]->entries.size() ? Lf[s]->entries[ef + 1].item0 : Lf[s]->total;
        r1 = int2{(int)(poff[s] + M.total + Lf[s]->entries[ef].item0), next - Lf[s]->entries[ef].item0};
      }
      ranges[((size_t)i * S + s) * 2] = r0;
      ranges[((size_t)i * S + s) * 2 + 1] = r1;
    }
  if (poff[S] > 0x7fffffffLL) return ss_set_error(RAFT_SS_EINVAL, "batch too large for the LL partials");
  if ((rc = ensure_buf(b, b->lay, sizeof(float2) * 2 * (size_t)b->lay_total)) ||
      (rc = ensure_buf(b, b->t0, sizeof(float) * (size_t)b->sino_total)) ||
      (rc = ensure_buf(b, b->t1, sizeof(float) * (size_t)b->sino_total)) ||
      (rc = ensure_buf(b, b->weights, sizeof(float) * (size_t)S * b->img_total)) ||
      (rc = ensure_buf(b, b->partials, sizeof(double) * (size_t)std::max<long long>(poff[S], 1))) ||
      (rc = ensure_buf(b, b->ll, sizeof(double) * (size_t)std::max(niter, 1) * ns)) ||
      (rc = ensure_buf(b, b->ranges, sizeof(int2) * ranges.size())) ||
      (rc = ensure_buf(b, b->slab, sizeof(float) * part_floats)))
    return rc;
  HIP_TRY(hipMemcpy(b->ranges.p, ranges.data(), sizeof(int2) * ranges.size(), hipMemcpyHostToDevice));
  const SSBatchSlice* ds = b->slices.as<SSBatchSlice>();
  float2* gA = b->lay.as<float2>();
  float2* gB = gA + b->lay_total;
  float* measT = b->t0.as<float>();
  float* rhoT = b->t1.as<float>();
  float* W = b->weights.as<float>();
  double* part = b->partials.as<double>();
  double* lld = b->ll.as<double>();
  const int2* rg = b->ranges.as<int2>();
  float* slab = b->slab.as<float>();

  // ---- W_b = BP_b(1) per slice and subset, computed once per (S, variants, buffers)
  if (b->weights_S != S || b->weights_generation != b->generation) {
    HIP_TRY(ss_launch_fill(rhoT, 1.0f, (size_t)b->sino_total, stream));
    for (int s = 0; s < S; ++s)
      for (size_t q = 0; q < vs.size(); ++q)
        HIP_TRY(ss_launch_batch_backprojection(Lbw[s * vs.size() + q]->dev(ds), vs[q], rhoT, SS_BP_PLAIN,
                                               W + (size_t)s * b->img_total, nullptr, nullptr, stream));
    b->weights_S = S;
    b->weights_generation = b->generation;
  }

  char key[160];
  snprintf(key, sizeof(key), "em|%p|%p|%d|%d", (const void*)d_sinos, (void*)d_x, niter, S);
  rc = run_body(b, flags & RAFT_SS_BATCH_GRAPH, key, stream, [&](hipStream_t st) -> int {
    HIP_TRY(ss_launch_batch_transpose(Lt->dev(ds), d_sinos, measT, 1, st));
    for (int it = 0; it < niter; ++it) {
      for (int s = 0; s < S; ++s) {
        HIP_TRY(ss_launch_batch_gather(Lg[s]->dev(ds), b->gather_tile, d_x, gA, gB, st));
        float* sp = Lr[s]->slabs ? slab : nullptr;
        HIP_TRY(ss_launch_batch_radon(Lr[s]->dev(ds), gA, gB, sp, SS_OUT_EM_RATIO, rhoT, measT, part + poff[s], st));
        if (sp)
          HIP_TRY(ss_launch_batch_radon_finish(Lz[s]->dev(ds), sp, SS_OUT_EM_RATIO, rhoT, measT, part + poff[s], st));
        HIP_TRY(ss_launch_batch_radon_fixup(Lf[s]->dev(ds), gA, gB, SS_OUT_EM_RATIO, rhoT, measT,
                                            part + poff[s] + Lm[s]->total, st));
        for (size_t q = 0; q < vs.size(); ++q)
          HIP_TRY(ss_launch_batch_backprojection(Lbu[s * vs.size() + q]->dev(ds), vs[q], rhoT, SS_BP_EM_UPDATE,
                                                 nullptr, d_x, W + (size_t)s * b->img_total, st));
      }
      HIP_TRY(ss_launch_batch_ll(rg, 2 * S, ns, part, lld + (size_t)it * ns, st));
    }
    return 0;
  });
  if (rc) return rc;
  if (ll_host && niter > 0) {
    HIP_TRY(hipMemcpyAsync(ll_host, lld, sizeof(double) * (size_t)niter * ns, hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
  }
  return 0;
}

// Autotuner for a batch: slices are grouped by image side n; for each group every
// variant that fits all its slices is timed as one batched backprojection of the
// group over angles [0, min(A_i, 360)), and the fastest is kept for the group.
extern "C" int raft_ss_batch_tune(raft_ss_batch* b, int reps, int flags, void* stream_) {
  (void)flags;
  int rc = check_batch(b);
  if (rc) return rc;
  std::lock_guard<std::mutex> lock(b->mu);
  hipStream_t stream = as_stream(stream_);
  std::map<int, std::vector<int>> groups;
  for (int i = 0; i < b->nslices; ++i) groups[b->plan(i).g.n].push_back(i);
  if ((rc = ensure_buf(b, b->t0, sizeof(float) * (size_t)b->sino_total)) ||
      (rc = ensure_buf(b, b->t1, sizeof(float) * (size_t)b->img_total)))
    return rc;
  HIP_TRY(ss_launch_fill(b->t0.as<float>(), 1.0f, (size_t)b->sino_total, stream));
  std::vector<int> chosen(b->variant);
  for (auto& kv : groups) {
    const std::vector<int>& members = kv.second;
    double best_t = 1e300;
    int best_v = -1;
    for (int v = 0; v < SS_BP_NVARIANTS; ++v) {
      bool ok = true;
      for (int i : members) ok = ok && b->plan(i).bp_angles_v[v] > 0;
      if (!ok) continue;
      // the group's slices with variant v, in a launch of their own
      std::vector<SSBatchSlice> h;
      BatchLaunch L;
      long long total = 0;
      for (int i : members) {
        SSBatchSlice sl;
        sl.p = b->plan(i).dev(v);
        sl.img_off = b->img_off[i];
        sl.sino_off = b->sino_off[i];
        sl.lay_off = b->lay_off[i];
        SSBatchEntry e{(int)h.size(), 0, std::min(sl.p.g.nangles, 360), (int)total};
        h.push_back(sl);
        L.entries.push_back(e);
        total += ss_bp_tiles(sl.p.g.n, v);
      }
      L.total = (int)total;
      L.grid = ss_batch_grid(SS_KIND_BP, v, SS_BP_PLAIN, L.total);
      Buf dsl;
      hipError_t e;
      dsl.ensure(sizeof(SSBatchSlice) * h.size(), &e);
      HIP_TRY(e);
      HIP_TRY(hipMemcpy(dsl.p, h.data(), sizeof(SSBatchSlice) * h.size(), hipMemcpyHostToDevice));
      HIP_TRY(hipMalloc(&L.d_entries, sizeof(SSBatchEntry) * L.entries.size()));
      HIP_TRY(hipMemcpy(L.d_entries, L.entries.data(), sizeof(SSBatchEntry) * L.entries.size(),
                        hipMemcpyHostToDevice));
      const SSBatchLaunch dl = L.dev(dsl.as<SSBatchSlice>());
      float* sT = b->t0.as<float>();
      float* img = b->t1.as<float>();
      double t = 0.0;
      auto run = [&]() -> int {
        HIP_TRY(ss_launch_batch_backprojection(dl, v, sT, SS_BP_PLAIN, img, nullptr, nullptr, stream));
        return 0;
      };
      if ((rc = time_best(run, reps, stream, &t))) return rc;
      if (t < best_t) {
        best_t = t;
        best_v = v;
      }
    }
    if (best_v < 0) return ss_set_error(RAFT_SS_EINTERNAL, "no variant fits the slices with n = %d", kv.first);
    for (int i : members) chosen[i] = best_v;
  }
  if (chosen != b->variant) {
    b->variant = chosen;
    variants_changed(b);
  }
  return 0;
}
