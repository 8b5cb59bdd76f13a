// mxr_tune.hip -- launch-shape autotuning of the fused find_HII_bubbles path and its JSON
// "wisdom" file.
//
// The reference keeps FFTW plans in a wisdom file named per grid size and rank count
// (src/core/reionization.c:726-758); here the tunable choices are the launch shapes of the
// fused path (mxr_tuning), measured on the device with find_HII_bubbles on the caller's
// inputs and stored per (device architecture, compute units, grid dimension, filter type,
// spin-temperature flag):
//
//   {"format": "mxr-wisdom-1",
//    "entries": [{"arch": "gfx950", "cus": 256, "N": 512, "filter": 1, "spin": 0,
//                 "zdirect_kmax": 16, "zdirect_rpb": 4, "live_tiles": 1,
//                 "col_tc_x": 16, "col_tc_y": 16, "ms": 41.1}, ...]}
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "mxr_internal.hpp"

namespace mxr {

// ---- the tuning record --------------------------------------------------------------------
static bool valid_tc(int tc) { return tc == 4 || tc == 8 || tc == 16 || tc == 32; }

int tuning_check(const mxr_tuning* t)
{
  if (t->zdirect_kmax < 0 || t->zdirect_kmax > 64 || t->zdirect_rpb < 1 || t->zdirect_rpb > 4096 ||
      (t->live_tiles != 0 && t->live_tiles != 1) || !valid_tc(t->col_tc_x) || !valid_tc(t->col_tc_y)) {
    set_error("tuning out of range: zdirect_kmax %d (0..64), zdirect_rpb %d (1..4096), live_tiles %d (0, 1), "
              "col_tc_x %d, col_tc_y %d (4, 8, 16, 32)",
              t->zdirect_kmax, t->zdirect_rpb, t->live_tiles, t->col_tc_x, t->col_tc_y);
    return MXR_ERR_INVALID_ARG;
  }
  return MXR_OK;
}

void tuning_get(const Ctx& c, mxr_tuning* t)
{
  t->zdirect_kmax = c.zdirect_kmax;
  t->zdirect_rpb = c.zdirect_rpb;
  t->live_tiles = c.live_tiles;
  t->col_tc_x = c.col_tc_x;
  t->col_tc_y = c.col_tc_y;
}

void tuning_set(Ctx& c, const mxr_tuning& t)
{
  c.zdirect_kmax = t.zdirect_kmax;
  c.zdirect_rpb = t.zdirect_rpb;
  c.live_tiles = t.live_tiles;
  c.col_tc_x = t.col_tc_x;
  c.col_tc_y = t.col_tc_y;
}

// ---- a small JSON reader and writer (objects, arrays, strings, numbers, literals) ----------
namespace {

struct JVal {
  enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
  double num = 0;
  std::string str;
  std::vector<JVal> arr;
  std::vector<std::pair<std::string, JVal>> obj;
  const JVal* get(const char* key) const
  {
    for (auto& kv : obj)
      if (kv.first == key)
        return &kv.second;
    return nullptr;
  }
};

struct JParser {
  const char* s;
  const char* end;
  bool ok = true;

  void ws()
  {
    while (s < end && (*s == ' ' || *s == '\t' || *s == '\n' || *s == '\r'))
      s++;
  }
  bool lit(const char* w)
  {
    const size_t n = std::strlen(w);
    if ((size_t)(end - s) >= n && std::strncmp(s, w, n) == 0) {
      s += n;
      return true;
    }
    return false;
  }
  bool string(std::string& out)
  {
    if (s >= end || *s != '"')
      return false;
    s++;
    while (s < end && *s != '"') {
      if (*s == '\\') {
        if (++s >= end)
          return false;
        const char e = *s;
        out += e == 'n' ? '\n' : e == 't' ? '\t' : e;  // the writer escapes only '"' and '\'
      } else {
        out += *s;
      }
      s++;
    }
    if (s >= end)
      return false;
    s++;
    return true;
  }
  bool value(JVal& v, int depth)
  {
    if (depth > 32)
      return false;
    ws();
    if (s >= end)
      return false;
    if (*s == '{') {
      v.kind = JVal::OBJ;
      s++;
      ws();
      if (s < end && *s == '}') {
        s++;
        return true;
      }
      for (;;) {
        ws();
        std::string k;
        if (!string(k))
          return false;
        ws();
        if (s >= end || *s != ':')
          return false;
        s++;
        JVal x;
        if (!value(x, depth + 1))
          return false;
        v.obj.emplace_back(std::move(k), std::move(x));
        ws();
        if (s < end && *s == ',') {
          s++;
          continue;
        }
        if (s < end && *s == '}') {
          s++;
          return true;
        }
        return false;
      }
    }
    if (*s == '[') {
      v.kind = JVal::ARR;
      s++;
      ws();
      if (s < end && *s == ']') {
        s++;
        return true;
      }
      for (;;) {
        JVal x;
        if (!value(x, depth + 1))
          return false;
        v.arr.push_back(std::move(x));
        ws();
        if (s < end && *s == ',') {
          s++;
          continue;
        }
        if (s < end && *s == ']') {
          s++;
          return true;
        }
        return false;
      }
    }
    if (*s == '"') {
      v.kind = JVal::STR;
      return string(v.str);
    }
    if (lit("true")) {
      v.kind = JVal::BOOL;
      v.num = 1;
      return true;
    }
    if (lit("false")) {
      v.kind = JVal::BOOL;
      return true;
    }
    if (lit("null"))
      return true;
    char* e = nullptr;
    std::string tmp(s, (size_t)std::min<long>(end - s, 64));
    v.num = std::strtod(tmp.c_str(), &e);
    if (e == tmp.c_str())
      return false;
    v.kind = JVal::NUM;
    s += e - tmp.c_str();
    return true;
  }
};

bool read_file(const char* path, std::string& out)
{
  FILE* f = std::fopen(path, "rb");
  if (!f)
    return false;
  char buf[4096];
  size_t n;
  while ((n = std::fread(buf, 1, sizeof buf, f)) > 0)
    out.append(buf, n);
  std::fclose(f);
  return true;
}

struct Entry {
  std::string arch;
  int cus = 0, N = 0, filter = 0, spin = 0;
  mxr_tuning t{};
  double ms = 0;
};

const char* kFields[] = {"zdirect_kmax", "zdirect_rpb", "live_tiles", "col_tc_x", "col_tc_y"};
int* field(mxr_tuning& t, int i)
{
  int* f[] = {&t.zdirect_kmax, &t.zdirect_rpb, &t.live_tiles, &t.col_tc_x, &t.col_tc_y};
  return f[i];
}

// entries of a wisdom file; a missing file is an empty list, a malformed one an error
int load_entries(const char* path, std::vector<Entry>& out)
{
  std::string text;
  if (!read_file(path, text))
    return errno == ENOENT ? MXR_OK : (set_error("cannot read wisdom file %s", path), MXR_ERR_INVALID_ARG);
  JParser p{text.data(), text.data() + text.size()};
  JVal root;
  if (!p.value(root, 0) || root.kind != JVal::OBJ) {
    set_error("wisdom file %s is not a JSON object", path);
    return MXR_ERR_INVALID_ARG;
  }
  const JVal* fmt = root.get("format");
  const JVal* ents = root.get("entries");
  if (!fmt || fmt->kind != JVal::STR || fmt->str != "mxr-wisdom-1" || !ents || ents->kind != JVal::ARR) {
    set_error("wisdom file %s: expected {\"format\": \"mxr-wisdom-1\", \"entries\": [...]}", path);
    return MXR_ERR_INVALID_ARG;
  }
  for (const JVal& e : ents->arr) {
    if (e.kind != JVal::OBJ)
      continue;
    Entry x;
    const JVal* a = e.get("arch");
    if (!a || a->kind != JVal::STR)
      continue;
    x.arch = a->str;
    bool ok = true;
    auto num = [&](const char* k, int& dst) {
      const JVal* v = e.get(k);
      if (!v || v->kind != JVal::NUM) {
        ok = false;
        return;
      }
      dst = (int)v->num;
    };
    num("cus", x.cus);
    num("N", x.N);
    num("filter", x.filter);
    num("spin", x.spin);
    for (int i = 0; i < 5; i++)
      num(kFields[i], *field(x.t, i));
    const JVal* ms = e.get("ms");
    x.ms = (ms && ms->kind == JVal::NUM) ? ms->num : 0.0;
    if (ok && tuning_check(&x.t) == MXR_OK)
      out.push_back(x);
  }
  return MXR_OK;
}

int write_entries(const char* path, const std::vector<Entry>& es)
{
  const std::string tmp = std::string(path) + ".tmp";
  FILE* f = std::fopen(tmp.c_str(), "wb");
  if (!f) {
    set_error("cannot write wisdom file %s", tmp.c_str());
    return MXR_ERR_INVALID_ARG;
  }
  std::fprintf(f, "{\"format\": \"mxr-wisdom-1\",\n \"entries\": [");
  for (size_t i = 0; i < es.size(); i++) {
    const Entry& e = es[i];
    std::string arch;
    for (char ch : e.arch) {
      if (ch == '"' || ch == '\\')
        arch += '\\';
      arch += ch;
    }
    std::fprintf(f, "%s\n  {\"arch\": \"%s\", \"cus\": %d, \"N\": %d, \"filter\": %d, \"spin\": %d", i ? "," : "",
                 arch.c_str(), e.cus, e.N, e.filter, e.spin);
    mxr_tuning t = e.t;
    for (int k = 0; k < 5; k++)
      std::fprintf(f, ", \"%s\": %d", kFields[k], *field(t, k));
    std::fprintf(f, ", \"ms\": %.4f}", e.ms);
  }
  std::fprintf(f, "\n ]}\n");
  const bool bad = std::ferror(f) != 0;
  if (std::fclose(f) != 0 || bad || std::rename(tmp.c_str(), path) != 0) {
    std::remove(tmp.c_str());
    set_error("cannot write wisdom file %s", path);
    return MXR_ERR_INVALID_ARG;
  }
  return MXR_OK;
}

bool same_key(const Entry& e, const char* arch, int cus, int N, int filter, int spin)
{
  return e.arch == arch && e.cus == cus && e.N == N && e.filter == filter && e.spin == spin;
}

__global__ void k_fill_f(float* __restrict__ p, float v, long n)
{
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    p[i] = v;
}

}  // namespace

int wisdom_lookup(const char* path, const char* arch, int cus, int N, int filter, int spin, mxr_tuning* t)
{
  std::vector<Entry> es;
  RC(load_entries(path, es));
  for (const Entry& e : es)
    if (same_key(e, arch, cus, N, filter, spin)) {
      *t = e.t;
      return MXR_OK;
    }
  set_error("wisdom file %s has no entry for %s, %d CUs, N = %d, filter %d, spin %d", path, arch, cus, N, filter,
            spin);
  return MXR_ERR_NO_WISDOM;
}

int wisdom_store(const char* path, const char* arch, int cus, int N, int filter, int spin, const mxr_tuning& t,
                 double ms)
{
  RC(tuning_check(&t));
  std::vector<Entry> es;
  RC(load_entries(path, es));
  Entry x;
  x.arch = arch;
  x.cus = cus;
  x.N = N;
  x.filter = filter;
  x.spin = spin;
  x.t = t;
  x.ms = ms;
  bool replaced = false;
  for (Entry& e : es)
    if (same_key(e, arch, cus, N, filter, spin)) {
      e = x;
      replaced = true;
    }
  if (!replaced)
    es.push_back(x);
  return write_entries(path, es);
}

// architecture name (up to the first ':') and compute-unit count of the context's device
int device_key(const Ctx& c, std::string& arch, int& cus)
{
  hipDeviceProp_t prop;
  MXR_HIP(hipGetDeviceProperties(&prop, c.device));
  arch = prop.gcnArchName;
  const size_t colon = arch.find(':');
  if (colon != std::string::npos)
    arch.resize(colon);
  cus = prop.multiProcessorCount;
  return MXR_OK;
}

// Coordinate search over the launch shapes: each field in turn takes every candidate value with
// the others at their best so far; a setting's time is the smaller of two calls after a warm-up.
// The calls run on scratch state grids, so the caller's state is not modified.
int autotune(Ctx* c, const mxr_hii_params* p, const mxr_hii_grids* g, mxr_tuning* best_out, double* best_ms)
{
  if (c->path != MXR_PATH_FUSED) {
    set_error("autotuning applies to the fused path");
    return MXR_ERR_UNSUPPORTED;
  }
  const int N = c->N;
  const long nreal = n_real(N), npad = 2 * n_complex(N);
  const bool bf16 = g->storage == MXR_STORAGE_BF16;
  const bool spin = p->Flag_IncludeSpinTemp != 0;
  std::vector<void*> owned;
  auto release = [&] {
    for (void* q : owned)
      (void)hipFree(q);
  };
  auto dev_alloc = [&](size_t bytes, void** q) -> int {
    const hipError_t e = hipMalloc(q, bytes);
    if (e != hipSuccess) {
      *q = nullptr;
      return hip_fail(e, "hipMalloc", __FILE__, __LINE__);
    }
    owned.push_back(*q);
    return MXR_OK;
  };
  mxr_hii_grids s = *g;
  int rc = MXR_OK;
  // inputs on the device
  if (g->mem == MXR_MEM_HOST) {
    const bool bh = p->Flag_BHFeedback != 0;
    const float* src[7] = {g->deltax,  g->stars, g->weighted_sfr, spin ? g->x_e_box : nullptr,
                           spin ? g->Tk_box : nullptr, bh ? g->effective_bhm : nullptr,
                           bh ? g->effective_bhar : nullptr};
    const float** dst[7] = {&s.deltax, &s.stars, &s.weighted_sfr, &s.x_e_box, &s.Tk_box, &s.effective_bhm,
                            &s.effective_bhar};
    for (int i = 0; i < 7 && rc == MXR_OK; i++) {
      if (!src[i])
        continue;
      const size_t bytes = i == 4 ? sizeof(float) * nreal
                                  : ((bf16 && (i == 1 || i == 2)) ? sizeof(uint16_t) : sizeof(float)) * npad;
      void* q = nullptr;
      rc = dev_alloc(bytes, &q);
      if (rc == MXR_OK) {
        const hipError_t e = hipMemcpy(q, src[i], bytes, hipMemcpyHostToDevice);
        if (e != hipSuccess)
          rc = hip_fail(e, "hipMemcpy", __FILE__, __LINE__);
      }
      *dst[i] = (const float*)q;
    }
  }
  // scratch state: xH, z_at_ionization, J_21, r_bubble, temp, and with recombinations N_rec
  // (padded), Gamma12, residual_xH, clumping_factor, t_resp
  const int nst = p->Flag_IncludeRecombinations ? 10 : 5;
  float* st[10] = {};
  for (int i = 0; i < nst && rc == MXR_OK; i++)
    rc = dev_alloc(sizeof(float) * (i == 5 ? npad : nreal), (void**)&st[i]);
  if (rc) {
    release();
    return rc;
  }
  s.mem = MXR_MEM_DEVICE;
  s.xH = st[0];
  s.z_at_ionization = st[1];
  s.J_21_at_ionization = st[2];
  s.r_bubble = st[3];
  s.temp_kinetic_all_gas = st[4];
  if (nst == 10) {
    s.N_rec = st[5];
    s.Gamma12 = st[6];
    s.residual_xH = st[7];
    s.clumping_factor = st[8];
    s.t_resp = st[9];
  }

  const bool saved_prof = c->prof;
  c->prof = false;
  mxr_tuning start;
  tuning_get(*c, &start);
  auto run_once = [&](double* ms) -> int {
    const float init[10] = {1.0f, -1.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    for (int i = 0; i < nst; i++)
      hipLaunchKernelGGL(k_fill_f, dim3(1024), dim3(256), 0, c->stream, st[i], init[i], i == 5 ? npad : nreal);
    MXR_HIP(hipGetLastError());
    mxr_hii_globals out;
    RC(run_find_hii_bubbles(c, p, &s, &out));
    *ms = out.device_ms;
    return MXR_OK;
  };
  auto measure = [&](const mxr_tuning& t, double* ms) -> int {
    tuning_set(*c, t);
    double a = 0, b = 0, w = 0;
    RC(run_once(&w));
    RC(run_once(&a));
    RC(run_once(&b));
    *ms = std::min(a, b);
    return MXR_OK;
  };
  mxr_tuning best = start;
  double bms = 0;
  rc = measure(best, &bms);
  struct Knob {
    int idx;
    std::vector<int> vals;
  };
  std::vector<Knob> knobs;
  const bool sharpk = p->ReionFilterType == 1;
  if (sharpk && N % 256 == 0) {
    knobs.push_back({0, {0, 8, 16, 24, 32, 40, 48, 56, 64}});
    knobs.push_back({1, {1, 2, 4, 8, 16, 32}});
  }
  knobs.push_back({4, {4, 8, 16, 32}});
  knobs.push_back({3, {4, 8, 16, 32}});
  if (sharpk)
    knobs.push_back({2, {0, 1}});
  for (const Knob& k : knobs) {
    if (rc)
      break;
    if (k.idx == 1 && best.zdirect_kmax == 0)
      continue;  // rows per wave only matter for the direct z pass
    for (int v : k.vals) {
      mxr_tuning t = best;
      if (*field(t, k.idx) == v)
        continue;
      *field(t, k.idx) = v;
      double ms = 0;
      rc = measure(t, &ms);
      if (rc)
        break;
      if (ms < bms) {
        bms = ms;
        best = t;
      }
    }
  }
  c->prof = saved_prof;
  tuning_set(*c, rc ? start : best);
  const hipError_t se = hipStreamSynchronize(c->stream);
  release();
  if (rc)
    return rc;
  if (se != hipSuccess)
    return hip_fail(se, "hipStreamSynchronize", __FILE__, __LINE__);
  *best_out = best;
  *best_ms = bms;
  return MXR_OK;
}

}  // namespace mxr
