// Launch-shape autotuner: times each candidate shape of a batch's model on the batch's own cells
// and keeps the fastest, with a persistent cache of choices keyed by device and workload class.
//
// Knobs (CEP_SHAPE_*): the register split of the one-lane-per-cell Rush-Larsen kernels (one wave
// per SIMD with all 512 registers, or two waves with 256 and spills), ORd 2011 cells per
// wavefront (64/32/16: the step split over 1, 2 or 4 wavefronts exchanging through LDS), the
// TT06 workgroup size and the dispatch order of the ORd-CiPA BDF kernel's 16-cell groups.  Which
// is fastest depends on the cell count against the 1024 SIMDs (one wave per SIMD only pays while
// the waves do not outnumber the SIMDs) and on how uneven the cells' beats are, so it is measured.
//
// A trial runs whole beats from the batch's current state on a bitwise snapshot of all its
// device memory, which is copied back after every trial and at the end: tuning changes nothing
// a later run computes.  Included by cep_gpu.hip after the batch and launch code.

namespace tune {

struct Shape {
  int v[CEP_NSHAPE];
};

// the candidates of the batch's model; empty when nothing is tunable (ORd 2011 Adams)
static std::vector<Shape> candidates(const cep_batch *b)
{
  std::vector<Shape> c;
  if (b->model == CEP_MODEL_ORD_CIPA) {
    c.push_back({{1, 16, 64, CEP_ORDER_STORAGE}});
    c.push_back({{1, 16, 64, CEP_ORDER_LONGEST_FIRST}});
  } else if (b->model == CEP_MODEL_ORD2011) {
    if (b->integrator == CEP_INTEGRATOR_ADAMS) return c;
    c.push_back({{2, 64, 64, 0}});
    c.push_back({{1, 64, 64, 0}});
    c.push_back({{1, 32, 128, 0}});
    c.push_back({{1, 16, 256, 0}});
  } else {
    for (int block : {256, 64})
      for (int wpe : {2, 1}) c.push_back({{wpe, 64, block, 0}});
  }
  return c;
}

// the shape the next launch uses, automatic choices resolved
static Shape current(const cep_batch *b)
{
  Shape s{{0, 0, 0, 0}};
  if (b->model == CEP_MODEL_ORD_CIPA) {
    s.v[CEP_SHAPE_WAVES_PER_EU] = 1;
    s.v[CEP_SHAPE_CELLS_PER_WAVE] = ordbdf::CPB;
    s.v[CEP_SHAPE_BLOCK] = 64;
    s.v[CEP_SHAPE_GROUP_ORDER] = bdf_group_order(b);
  } else if (b->model == CEP_MODEL_ORD2011) {
    const int w = b->integrator == CEP_INTEGRATOR_ADAMS ? 1 : ord11_waves(b);
    s.v[CEP_SHAPE_WAVES_PER_EU] = w == 1 && b->integrator != CEP_INTEGRATOR_ADAMS ? rl_waves_per_eu(b) : 1;
    s.v[CEP_SHAPE_CELLS_PER_WAVE] = 64 / w;
    s.v[CEP_SHAPE_BLOCK] = 64 * w;
  } else {
    s.v[CEP_SHAPE_WAVES_PER_EU] = rl_waves_per_eu(b);
    s.v[CEP_SHAPE_CELLS_PER_WAVE] = 64;
    s.v[CEP_SHAPE_BLOCK] = tt_block(b);
  }
  return s;
}

// sets one knob; value 0 returns it to automatic.  Returns false for a value the model has no
// kernel for.
static bool set(cep_batch *b, int knob, int value)
{
  switch (knob) {
  case CEP_SHAPE_WAVES_PER_EU:
    if (b->model == CEP_MODEL_ORD_CIPA) return value == 0 || value == 1;
    if (value != 0 && value != 1 && value != 2) return false;
    b->shape_wpe = value;
    return true;
  case CEP_SHAPE_CELLS_PER_WAVE:
    if (b->model == CEP_MODEL_ORD_CIPA) return value == 0 || value == ordbdf::CPB;
    if (b->model == CEP_MODEL_TT06) return value == 0 || value == 64;
    if (value != 0 && value != 64 && value != 32 && value != 16) return false;
    b->o11_waves = value == 0 ? 0 : 64 / value;
    return true;
  case CEP_SHAPE_BLOCK:
    if (b->model == CEP_MODEL_TT06) {
      if (value != 0 && value != 64 && value != 128 && value != 256) return false;
      b->shape_block = value;
      return true;
    }
    return value == 0 || value == current(b).v[CEP_SHAPE_BLOCK];
  case CEP_SHAPE_GROUP_ORDER:
    if (b->model != CEP_MODEL_ORD_CIPA) return value == 0;
    if (value != 0 && value != CEP_ORDER_STORAGE && value != CEP_ORDER_LONGEST_FIRST) return false;
    b->shape_order = value;
    return true;
  }
  return false;
}

static void apply(cep_batch *b, const Shape &s, bool only_automatic)
{
  if (b->model == CEP_MODEL_ORD_CIPA) {
    if (!only_automatic || b->shape_order == 0) b->shape_order = s.v[CEP_SHAPE_GROUP_ORDER];
  } else if (b->model == CEP_MODEL_ORD2011) {
    if (!only_automatic || b->o11_waves == 0) b->o11_waves = 64 / s.v[CEP_SHAPE_CELLS_PER_WAVE];
    if (!only_automatic || b->shape_wpe == 0) b->shape_wpe = s.v[CEP_SHAPE_WAVES_PER_EU];
  } else {
    if (!only_automatic || b->shape_wpe == 0) b->shape_wpe = s.v[CEP_SHAPE_WAVES_PER_EU];
    if (!only_automatic || b->shape_block == 0) b->shape_block = s.v[CEP_SHAPE_BLOCK];
  }
}

// the cache key: device architecture and CU count, model, precision, integrator and the cell
// count rounded up to a power of two
static std::string key(const cep_batch *b)
{
  int dev = 0;
  hipDeviceProp_t p{};
  std::string arch = "unknown";
  if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&p, dev) == hipSuccess) {
    arch = p.gcnArchName;
    const size_t colon = arch.find(':');
    if (colon != std::string::npos) arch.resize(colon);
  }
  long long cells = 1;
  while (cells < b->n) cells <<= 1;
  char buf[200];
  snprintf(buf, sizeof buf, "%s cus=%d model=%d prec=%d integrator=%d cells<=%lld", arch.c_str(), device_cus(),
           b->model, b->prec, b->integrator, cells);
  return buf;
}

static std::vector<std::string> read_lines(const char *path)
{
  std::vector<std::string> lines;
  FILE *f = fopen(path, "r");
  if (!f) return lines;
  char buf[512];
  while (fgets(buf, sizeof buf, f)) {
    std::string l(buf);
    while (!l.empty() && (l.back() == '\n' || l.back() == '\r')) l.pop_back();
    if (!l.empty()) lines.push_back(l);
  }
  fclose(f);
  return lines;
}

// one line per key: "<key> | wpe=W cpw=C block=B order=O | ms=T trial_beats=K"
static bool parse(const std::string &line, const std::string &k, Shape &s)
{
  if (line.compare(0, k.size() + 3, k + " | ") != 0) return false;
  return sscanf(line.c_str() + k.size() + 3, "wpe=%d cpw=%d block=%d order=%d", &s.v[0], &s.v[1], &s.v[2],
                &s.v[3]) == 4;
}

static int store(const char *path, const std::string &k, const Shape &s, double ms, int trial_beats)
{
  std::vector<std::string> keep;
  for (const std::string &l : read_lines(path)) {
    Shape t;
    if (!parse(l, k, t)) keep.push_back(l);
  }
  char buf[512];
  snprintf(buf, sizeof buf, "%s | wpe=%d cpw=%d block=%d order=%d | ms=%.3f trial_beats=%d", k.c_str(), s.v[0],
           s.v[1], s.v[2], s.v[3], ms, trial_beats);
  keep.push_back(buf);
  const std::string tmp = std::string(path) + ".tmp";
  FILE *f = fopen(tmp.c_str(), "w");
  if (!f) return -1;
  for (const std::string &l : keep) fprintf(f, "%s\n", l.c_str());
  if (fclose(f) != 0) return -1;
  return rename(tmp.c_str(), path) == 0 ? 0 : -1;   // whole file replaced at once
}

// a bitwise copy of every device block of the batch
struct Snapshot {
  std::vector<void *> copy;
  void release()
  {
    for (void *p : copy)
      if (p) (void)hipFree(p);
    copy.clear();
  }
};

static int take(cep_batch *b, Snapshot &s)
{
  for (const auto &a : b->allocs) {
    void *p = nullptr;
    const hipError_t e = hipMalloc(&p, a.second);
    if (e != hipSuccess) return fail(b, CEP_E_HIP, std::string("tuning snapshot: hipMalloc: ") + hipGetErrorString(e));
    s.copy.push_back(p);
    HIPCHK(b, hipMemcpyAsync(p, a.first, a.second, hipMemcpyDeviceToDevice, b->stream));
  }
  return 0;
}

static int put_back(cep_batch *b, const Snapshot &s)
{
  for (size_t i = 0; i < s.copy.size(); ++i)
    HIPCHK(b, hipMemcpyAsync(b->allocs[i].first, s.copy[i], b->allocs[i].second, hipMemcpyDeviceToDevice, b->stream));
  return 0;
}

}  // namespace tune

extern "C" {

int cep_set_launch_shape(cep_batch *b, int knob, int value)
{
  if (!b) return CEP_E_ARG;
  if (knob < 0 || knob >= CEP_NSHAPE) return fail(b, CEP_E_ARG, "unknown launch-shape knob");
  if (!tune::set(b, knob, value)) return fail(b, CEP_E_ARG, "no kernel of this model has that launch shape");
  return 0;
}

int cep_get_launch_shape(cep_batch *b, int knob, int *value)
{
  if (!b || !value) return CEP_E_ARG;
  if (knob < 0 || knob >= CEP_NSHAPE) return fail(b, CEP_E_ARG, "unknown launch-shape knob");
  *value = tune::current(b).v[knob];
  return 0;
}

int cep_tune_lookup(cep_batch *b, const char *cache_path)
{
  if (!b) return CEP_E_ARG;
  if (!cache_path || !*cache_path) return 0;
  const std::string k = tune::key(b);
  const std::vector<tune::Shape> cand = tune::candidates(b);
  for (const std::string &l : tune::read_lines(cache_path)) {
    tune::Shape s;
    if (!tune::parse(l, k, s)) continue;
    for (const tune::Shape &c : cand)
      if (memcmp(c.v, s.v, sizeof s.v) == 0) {
        tune::apply(b, s, true);
        return 1;
      }
  }
  return 0;
}

int cep_autotune(cep_batch *b, int trial_beats, const char *cache_path, double *report, int max_candidates,
                 int *ncandidates)
{
  if (!b || trial_beats < 1) return fail(b, CEP_E_ARG, "trial_beats >= 1");
  if (b->cap == 0) return fail(b, CEP_E_STATE, "cep_init_consts not called");
  const std::vector<tune::Shape> cand = tune::candidates(b);
  if (cand.empty()) return fail(b, CEP_E_ARG, "nothing to tune for this model and integrator");
  if (int rc = cep_sync(b)) return rc;
  if (int rc = alloc_trace_buffers(b)) return rc;    // part of the snapshot, not allocated by a trial
  const int shape_wpe = b->shape_wpe, shape_block = b->shape_block, shape_order = b->shape_order;
  const int o11_waves = b->o11_waves, beats_done = b->beats_done, klaunches = b->klaunches;
  const double ktime_total = b->ktime_total;
  tune::Snapshot snap;
  int rc = tune::take(b, snap);
  std::vector<double> ms(cand.size(), 0.0);
  for (size_t i = 0; i < cand.size() && rc == 0; ++i) {
    if (i > 0) rc = tune::put_back(b, snap);
    if (rc) break;
    b->beats_done = beats_done;
    tune::apply(b, cand[i], false);
    rc = cep_run_beats_async(b, trial_beats);
    if (rc) break;
    if (hipStreamSynchronize(b->stream) != hipSuccess) {
      rc = fail(b, CEP_E_HIP, "tuning trial failed");
      break;
    }
    float t = 0.f;
    if (hipEventElapsedTime(&t, b->ev0, b->ev1) != hipSuccess) {
      rc = fail(b, CEP_E_HIP, "tuning trial timing");
      break;
    }
    b->pending = false;
    ms[i] = t;
  }
  if (!snap.copy.empty() && tune::put_back(b, snap) == 0) (void)hipStreamSynchronize(b->stream);
  snap.release();
  b->shape_wpe = shape_wpe; b->shape_block = shape_block; b->shape_order = shape_order;
  b->o11_waves = o11_waves; b->beats_done = beats_done;
  b->klaunches = klaunches; b->ktime_total = ktime_total; b->pending = false;
  if (rc) return rc;
  size_t best = 0;
  for (size_t i = 1; i < cand.size(); ++i)
    if (ms[i] < ms[best]) best = i;
  tune::apply(b, cand[best], false);
  if (ncandidates) *ncandidates = (int)cand.size();
  if (report)
    for (size_t i = 0; i < cand.size() && (int)i < max_candidates; ++i) {
      for (int k = 0; k < CEP_NSHAPE; ++k) report[i * (CEP_NSHAPE + 1) + k] = cand[i].v[k];
      report[i * (CEP_NSHAPE + 1) + CEP_NSHAPE] = ms[i];
    }
  if (cache_path && *cache_path && tune::store(cache_path, tune::key(b), cand[best], ms[best], trial_beats) != 0)
    return fail(b, CEP_E_ARG, std::string("cannot write the tuning cache ") + cache_path);
  return (int)best;
}

}  // extern "C"
