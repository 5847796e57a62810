// lufqhe_table.cpp -- the autotuner's shape table, read from C.
//
// lion-codes_amd/autotune.py writes one JSON file per device kind:
//   {"device": ..., "entries": [{"policy": "crout_ref", "n": 32, "shape": 0,
//     "lanes": 32, "lane_cols": 2, "waves": 1, ...}, ...]}
// lufqhe_load_shape_table selects each entry's shape for the fused
// arithmetic, so C and C++ callers (LUCoreGPU::CALL included) run the tuned
// kernels without the Python package.  An entry applies only while the
// compiled candidate at that index still has the recorded geometry.
#include <dlfcn.h>

#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/lufqhe.h"

namespace {

bool read_file(const std::string& path, std::string& out) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return false;
  char buf[1 << 16];
  size_t n;
  while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) out.append(buf, n);
  std::fclose(f);
  return true;
}

// Value of "key" inside one flat JSON object: a string (quotes stripped) or
// a number, as text; empty if absent.
std::string field(const std::string& obj, const char* key) {
  const std::string k = std::string("\"") + key + "\"";
  size_t p = obj.find(k);
  if (p == std::string::npos) return {};
  p = obj.find(':', p + k.size());
  if (p == std::string::npos) return {};
  ++p;
  while (p < obj.size() && (obj[p] == ' ' || obj[p] == '\n' || obj[p] == '\t')) ++p;
  if (p < obj.size() && obj[p] == '"') {
    const size_t e = obj.find('"', p + 1);
    return e == std::string::npos ? std::string() : obj.substr(p + 1, e - p - 1);
  }
  size_t e = p;
  while (e < obj.size() && (std::isdigit((unsigned char)obj[e]) || obj[e] == '-' ||
                            obj[e] == '.' || obj[e] == 'e' || obj[e] == 'E' || obj[e] == '+'))
    ++e;
  return obj.substr(p, e - p);
}

std::string default_path() {
  if (const char* env = std::getenv("LUFQHE_SHAPE_TABLE")) return env;
  Dl_info info;
  if (dladdr(reinterpret_cast<void*>(&lufqhe_load_shape_table), &info) && info.dli_fname) {
    std::string lib = info.dli_fname;
    const size_t slash = lib.rfind('/');
    return (slash == std::string::npos ? std::string(".") : lib.substr(0, slash)) +
           "/shapes_gfx950.json";
  }
  return "shapes_gfx950.json";
}

}  // namespace

extern "C" int32_t lufqhe_load_shape_table(const char* path) {
  std::string text;
  if (!read_file(path ? std::string(path) : default_path(), text)) return -1;
  size_t p = text.find("\"entries\"");
  if (p == std::string::npos) return -1;
  p = text.find('[', p);
  if (p == std::string::npos) return -1;
  int32_t applied = 0;
  int depth = 0;
  size_t start = 0;
  for (size_t i = p + 1; i < text.size(); ++i) {
    const char c = text[i];
    if (c == '{') {
      if (depth++ == 0) start = i;
    } else if (c == '}') {
      if (--depth == 0) {
        const std::string obj = text.substr(start, i - start + 1);
        const std::string pol = field(obj, "policy");
        const int policy = pol == "crout_ref" ? LUFQHE_POLICY_CROUT_REF
                           : pol == "abs"     ? LUFQHE_POLICY_ABS
                                              : -1;
        const std::string sn = field(obj, "n"), ss = field(obj, "shape");
        const std::string sl = field(obj, "lanes"), sc = field(obj, "lane_cols");
        const std::string sw = field(obj, "waves");
        if (policy < 0 || sn.empty() || ss.empty() || sl.empty() || sc.empty()) continue;
        const int n = std::atoi(sn.c_str()), shape = std::atoi(ss.c_str());
        const int want_w = sw.empty() ? 1 : std::atoi(sw.c_str());
        int32_t lanes = 0, cols = 0, waves = 0;
        if (lufqhe_shape_info(n, shape, &lanes, &cols, &waves) != LUFQHE_SUCCESS) continue;
        if (lanes != std::atoi(sl.c_str()) || cols != std::atoi(sc.c_str()) || waves != want_w)
          continue;
        if (lufqhe_set_shape(policy, LUFQHE_ARITH_FUSED, n, shape) == LUFQHE_SUCCESS) ++applied;
      }
    } else if (c == ']' && depth == 0) {
      break;
    }
  }
  return applied;
}
