// tune.cpp -- autotuner of the batched sampler's launch shape.
//
// The reference schedules every row loop the same way (OpenMP
// "dynamic, 2", latentprior.cpp:73).  The batched sampler's schedule has four
// knobs: the segment length that splits heavy rows, and the work-group caps
// of the light-row, heavy-row and segment kernels.  Their best values depend
// on the latent size and on the row-degree distribution of each mode, so
// they are timed on the device: for each mode, every candidate shape runs a
// few half-sweeps between HIP events and the fastest is kept (the default
// stays unless a candidate beats it by more than 1%).  The factors of every
// mode are saved before and restored after the timing runs, so tuning leaves
// the chain where it was.
//
// Choices are recorded per (GPU architecture, padded latent size, number of
// other modes, rows, nonzeros, histogram of log2 row degrees) in a text table
// and looked up at initialisation: first in the file named by
// MACAU_AMD_TUNE_CACHE, then in tuning.txt beside the library
// (MACAU_AMD_TUNE_DISABLE=1 ignores both and keeps the default shapes).
#include <dlfcn.h>

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>

#include "model.h"

namespace mgpu {

namespace {

std::string gpu_arch() {
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return "unknown";
  std::string a = prop.gcnArchName;
  const size_t c = a.find(':');
  return c == std::string::npos ? a : a.substr(0, c);
}

std::string library_dir() {
  Dl_info info;
  if (dladdr(reinterpret_cast<void*>(&gpu_arch), &info) && info.dli_fname) {
    std::string p = info.dli_fname;
    const size_t s = p.rfind('/');
    return s == std::string::npos ? std::string(".") : p.substr(0, s);
  }
  return ".";
}

struct TuneEntry {
  std::string key;
  TuneChoice c;
};

std::vector<TuneEntry> read_table(const std::string& path) {
  std::vector<TuneEntry> out;
  std::ifstream f(path.c_str());
  std::string line;
  while (std::getline(f, line)) {
    if (line.empty() || line[0] == '#') continue;
    std::istringstream in(line);
    std::string arch, kp, no, n, nnz, h;
    TuneEntry e;
    if (!(in >> arch >> kp >> no >> n >> nnz >> h >> e.c.seg_len >> e.c.light_cap >> e.c.heavy_cap >> e.c.seg_cap))
      continue;
    in >> e.c.default_ms >> e.c.ms;
    e.key = arch + " " + kp + " " + no + " " + n + " " + nnz + " " + h;
    e.c.tuned = true;
    out.push_back(e);
  }
  return out;
}

}  // namespace

std::string Model::tune_key(int mode) const {
  // histogram of floor(log2(degree + 1)), hashed (FNV-1a 64)
  const std::vector<int>& rp = rows[mode].row_ptr;
  std::vector<long long> hist(33, 0);
  for (size_t n = 0; n + 1 < rp.size(); n++) {
    unsigned d = (unsigned)(rp[n + 1] - rp[n]) + 1u;
    int b = 0;
    while (d >>= 1) b++;
    hist[b]++;
  }
  uint64_t h = 1469598103934665603ull;
  for (long long v : hist)
    for (int k = 0; k < 8; k++) {
      h ^= (uint64_t)((v >> (8 * k)) & 0xff);
      h *= 1099511628211ull;
    }
  char buf[256];
  snprintf(buf, sizeof(buf), "%s %d %d %d %lld %016llx", gpu_arch().c_str(), KP, nmodes - 1, dims[mode],
           (long long)rp.back(), (unsigned long long)h);
  return buf;
}

void Model::apply_cached_tuning() {
  tune.assign(nmodes, TuneChoice());
  if (sampler != SAMPLER_BATCHED) return;
  if (const char* off = std::getenv("MACAU_AMD_TUNE_DISABLE"))
    if (off[0] && off[0] != '0') return;  // the default shapes
  std::vector<TuneEntry> table;
  if (const char* p = std::getenv("MACAU_AMD_TUNE_CACHE")) table = read_table(p);
  for (auto& e : read_table(library_dir() + "/tuning.txt")) table.push_back(e);
  for (int m = 0; m < nmodes; m++) {
    const std::string k = tune_key(m);
    for (const auto& e : table)
      if (e.key == k) {
        tune[m] = e.c;
        break;
      }
  }
}

void Model::apply_plan_shape(int m) {
  plans[m].light_cap = tune[m].light_cap;
  plans[m].heavy_cap = tune[m].heavy_cap;
  plans[m].seg_cap = tune[m].seg_cap;
}

void Model::update_sse_slots() {
  if (!fused_sse) return;
  const int last = nmodes - 1;
  sse_partial_n = batched_light_grid(plans[last]) + batched_heavy_grid(plans[last]);
}

std::vector<std::pair<double, double>> Model::autotune(int reps, const std::string& record_path) {
  if (!initialized) throw std::runtime_error("model not initialized");
  std::vector<std::pair<double, double>> out(nmodes, {0.0, 0.0});
  if (sampler != SAMPLER_BATCHED) return out;
  if (reps < 1) reps = 1;
  // the iteration graphs hold the old launch shapes
  drop_graphs();
  launch_pending_prior();
  HIPCHK(hipStreamSynchronize(stream));
  std::vector<DevBuf> saved(nmodes);
  for (int m = 0; m < nmodes; m++) {
    saved[m].alloc(d_U[m].bytes());
    HIPCHK(hipMemcpyAsync(saved[m].as<void>(), d_U[m].as<void>(), d_U[m].bytes(), hipMemcpyDeviceToDevice, stream));
  }
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  auto time_mode = [&](int m, int r) {
    priors[m]->sample_latents(*this, m);  // warm
    HIPCHK(hipEventRecord(e0, stream));
    for (int q = 0; q < r; q++) priors[m]->sample_latents(*this, m);
    HIPCHK(hipEventRecord(e1, stream));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    return (double)ms / r;
  };
  const int seg_lens[] = {1024, 2048, 4096};
  const int light_caps[] = {2048, 4096, 8192, 16384};
  const int heavy_caps[] = {1024, 2048, 4096};
  const int seg_caps[] = {4096, 16384, 65536};
  std::vector<std::string> records;
  for (int m = 0; m < nmodes; m++) {
    const TuneChoice def;
    TuneChoice best = def;
    // the default shape, measured with the same repetitions as the candidates
    tune[m] = def;
    plans[m] = build_batch_plan(rows[m].row_ptr, KP, plan_bufs[m], stream, def.seg_len);
    apply_plan_shape(m);
    double best_ms = time_mode(m, reps);
    for (int sl : seg_lens) {
      plans[m] = build_batch_plan(rows[m].row_ptr, KP, plan_bufs[m], stream, sl);
      for (int lc : light_caps)
        for (int hc : heavy_caps)
          for (int sc : seg_caps) {
            TuneChoice c;
            c.seg_len = sl;
            c.light_cap = lc;
            c.heavy_cap = hc;
            c.seg_cap = sc;
            tune[m] = c;
            apply_plan_shape(m);
            const double ms = time_mode(m, reps);
            if (ms < best_ms) {
              best_ms = ms;
              best = c;
            }
          }
    }
    // confirm against the default with longer runs; keep the default unless
    // the candidate is more than 1% faster
    const int confirm = 4 * reps;
    tune[m] = def;
    plans[m] = build_batch_plan(rows[m].row_ptr, KP, plan_bufs[m], stream, def.seg_len);
    apply_plan_shape(m);
    const double d_ms = time_mode(m, confirm);
    tune[m] = best;
    plans[m] = build_batch_plan(rows[m].row_ptr, KP, plan_bufs[m], stream, best.seg_len);
    apply_plan_shape(m);
    const double b_ms = time_mode(m, confirm);
    if (!(b_ms < 0.99 * d_ms)) {
      best = def;
      tune[m] = def;
      plans[m] = build_batch_plan(rows[m].row_ptr, KP, plan_bufs[m], stream, def.seg_len);
      apply_plan_shape(m);
    }
    tune[m].tuned = true;
    tune[m].default_ms = d_ms;
    tune[m].ms = std::min(d_ms, b_ms);
    out[m] = {d_ms, tune[m].ms};
    char buf[512];
    snprintf(buf, sizeof(buf), "%s %d %d %d %d %.5f %.5f", tune_key(m).c_str(), tune[m].seg_len, tune[m].light_cap,
             tune[m].heavy_cap, tune[m].seg_cap, d_ms, tune[m].ms);
    records.push_back(buf);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  for (int m = 0; m < nmodes; m++)
    HIPCHK(hipMemcpyAsync(d_U[m].as<void>(), saved[m].as<void>(), d_U[m].bytes(), hipMemcpyDeviceToDevice, stream));
  // timing runs from the initial state may flag rows (the probit
  // augmentation has not run yet): the chain starts with a clear flag
  HIPCHK(hipMemsetAsync(d_fail.as<void>(), 0, sizeof(int), stream));
  HIPCHK(hipStreamSynchronize(stream));
  update_sse_slots();
  if (!record_path.empty()) {
    std::ofstream f(record_path.c_str(), std::ios::app);
    if (!f) throw std::runtime_error("cannot append the tuning record to " + record_path);
    for (auto& r : records) f << r << "\n";
  }
  return out;
}

}  // namespace mgpu
