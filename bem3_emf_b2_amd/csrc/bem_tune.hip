// bem_tune.hip -- the tuning table: schedule and tile choices of the
// library that leave results unchanged up to the FP32 rounding of the
// mixed form's refined products (which product kernel takes a shape, how
// many device streams the per-object work spreads over, whether the
// objects' assembly runs as one batch).  The entries start at the defaults
// below, measured on the 16-object scene; bh_init then loads the table
// named by BEM3_HIP_TUNE, or else tune_gfx950.txt beside the library when
// it exists (written by the autotuner, bem3_emf_b2_amd/autotune.py, from
// measurements on the device), unless bh_tune_set or bh_tune_load ran
// first; the environment switches BEM3_HIP_LANES, BEM3_HIP_ASM_BATCH and
// BEM3_HIP_GEMM_BF6 override single entries.
#include <dlfcn.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "bem_obj.h"

namespace {

struct TuneEntry {
  const char *name;
  int lo, hi, def;
  const char *env;
};

// lanes: 2 streams 6.9 s, 4 6.35 s, 6 6.1 s, 8 6.5 s per scene configuration;
// the product thresholds: the 128x128-tile kernels overtake the 64x64 ones
// from about 160 tiles (k >= 32), with k >= 1024 from about 64 tiles
const TuneEntry k_tab[BH_TUNE_COUNT] = {
    {"lanes", 1, 8, 6, "BEM3_HIP_LANES"},
    {"asm_batch", 0, 1, 1, "BEM3_HIP_ASM_BATCH"},
    {"gemm_tiles_fp64", 1, 1 << 20, 160, nullptr},
    {"gemm_tiles_fp32", 1, 1 << 20, 160, nullptr},
    {"gemm_longk_tiles", 1, 1 << 20, 64, nullptr},
    {"gemm_longk_k", 32, 1 << 20, 1024, nullptr},
    {"gemm_bf6", 0, 1, 1, "BEM3_HIP_GEMM_BF6"},
};

int g_val[BH_TUNE_COUNT];
bool g_ready = false;
std::string g_src = "defaults";

void tune_defaults()
{
  for (int i = 0; i < BH_TUNE_COUNT; i++) g_val[i] = k_tab[i].def;
  g_src = "defaults";
  g_ready = true;
}

// "name value" per line, '#' starts a comment; 0 or a message
int tune_parse(FILE *f, const char *path, int *val)
{
  char line[512];
  int ln = 0;
  while (fgets(line, sizeof line, f)) {
    ln++;
    char *h = strchr(line, '#');
    if (h) *h = 0;
    char name[128];
    long v = 0;
    char extra[8];
    const int n = sscanf(line, "%127s %ld %7s", name, &v, extra);
    if (n <= 0) continue;
    if (n != 2) return bem::fail(BH_EARG, "%s:%d: expected \"name value\"", path, ln);
    int k = 0;
    while (k < BH_TUNE_COUNT && strcmp(k_tab[k].name, name) != 0) k++;
    if (k == BH_TUNE_COUNT) return bem::fail(BH_EARG, "%s:%d: unknown tuning entry \"%s\"", path, ln, name);
    if (v < k_tab[k].lo || v > k_tab[k].hi)
      return bem::fail(BH_EARG, "%s:%d: %s = %ld outside %d..%d", path, ln, name, v, k_tab[k].lo, k_tab[k].hi);
    val[k] = (int)v;
  }
  return 0;
}

void tune_env()
{
  for (int i = 0; i < BH_TUNE_COUNT; i++) {
    const char *e = k_tab[i].env ? getenv(k_tab[i].env) : nullptr;
    if (!e || !*e) continue;
    const long v = strtol(e, nullptr, 10);
    g_val[i] = (int)(v < k_tab[i].lo ? k_tab[i].lo : v > k_tab[i].hi ? k_tab[i].hi : v);
  }
}

}  // namespace

namespace bem {

int tune_value(int key) { return (key >= 0 && key < BH_TUNE_COUNT) ? (g_ready ? g_val[key] : k_tab[key].def) : 0; }

// the table of bh_init: BEM3_HIP_TUNE, else tune_gfx950.txt beside the
// library, then the environment switches; a table that bh_tune_set or
// bh_tune_load filled before stays as it is
int tune_init()
{
  if (g_ready) return 0;
  tune_defaults();
  const char *p = getenv("BEM3_HIP_TUNE");
  std::string path;
  if (p && *p) {
    path = p;
  } else {
    Dl_info di;
    if (dladdr((const void *)&tune_init, &di) && di.dli_fname) {
      std::string lib = di.dli_fname;
      const size_t s = lib.rfind('/');
      path = (s == std::string::npos ? std::string(".") : lib.substr(0, s)) + "/tune_gfx950.txt";
      FILE *t = fopen(path.c_str(), "r");
      if (!t) path.clear(); else fclose(t);
    }
  }
  if (!path.empty()) {
    int rc = bh_tune_load(path.c_str());
    if (rc) { g_ready = false; return rc; }
  }
  tune_env();
  return 0;
}

}  // namespace bem

extern "C" int bh_tune_load(const char *path)
{
  API_LOCK;
  if (!path) return bem::fail(BH_EARG, "bh_tune_load: no path");
  FILE *f = fopen(path, "r");
  if (!f) return bem::fail(BH_EARG, "tuning table %s: cannot open", path);
  int val[BH_TUNE_COUNT];
  for (int i = 0; i < BH_TUNE_COUNT; i++) val[i] = k_tab[i].def;
  const int rc = tune_parse(f, path, val);
  fclose(f);
  if (rc) return rc;
  memcpy(g_val, val, sizeof val);
  g_src = path;
  g_ready = true;
  return 0;
}

extern "C" int bh_tune_set(int key, int value)
{
  API_LOCK;
  if (key < 0 || key >= BH_TUNE_COUNT) return bem::fail(BH_EARG, "tuning entry %d (0..%d)", key, BH_TUNE_COUNT - 1);
  if (value < k_tab[key].lo || value > k_tab[key].hi)
    return bem::fail(BH_EARG, "%s = %d outside %d..%d", k_tab[key].name, value, k_tab[key].lo, k_tab[key].hi);
  if (!g_ready) tune_defaults();
  g_val[key] = value;
  return 0;
}

extern "C" int bh_tune_get(int key, int *value)
{
  API_LOCK;
  if (key < 0 || key >= BH_TUNE_COUNT || !value) return bem::fail(BH_EARG, "tuning entry %d (0..%d)", key, BH_TUNE_COUNT - 1);
  *value = bem::tune_value(key);
  return 0;
}

extern "C" const char *bh_tune_name(int key)
{
  return (key >= 0 && key < BH_TUNE_COUNT) ? k_tab[key].name : nullptr;
}

extern "C" const char *bh_tune_source(void)
{
  API_LOCK;
  static thread_local std::string s;
  s = g_src;
  return s.c_str();
}
