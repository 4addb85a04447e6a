// pa_tune.hip -- the tuning state's C ABI and its persistence: chosen sort
// shapes per target and size class, saved to and loaded from a JSON cache
// keyed by the GPU architecture and the HIP runtime version.
//
// No reference counterpart: the reference fixes its radix and work quanta at
// compile time (RadixSortLsdParallel.h:215-216,368-379) and lists persisting
// tuned settings as future work (TODO.txt:43).
//
// Cache layout (shapes are stored by name, so a rebuilt library with a
// different candidate order still reads it; unknown names are skipped):
//   {"format": "pa-tune-1",
//    "devices": {"<arch>|hip <version>": {"u32": [name|null x 4],
//                                         "u64": [name|null x 4],
//                                         "seg_u32": [name|null]}}}
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "pa_internal.h"

namespace pa {

// ---------------------------------------------------------------------------
// A small JSON reader and writer for the cache (objects, arrays, strings,
// null; numbers and booleans are read and kept as text).
// ---------------------------------------------------------------------------
struct JVal {
  enum Type { Null, Str, Num, Obj, Arr } t = Null;
  std::string s;
  std::vector<std::pair<std::string, JVal>> obj;
  std::vector<JVal> arr;
  const JVal* get(const std::string& k) const {
    for (auto& kv : obj)
      if (kv.first == k) return &kv.second;
    return nullptr;
  }
};

struct JParser {
  const std::string& x;
  size_t i = 0;
  bool ok = true;
  explicit JParser(const std::string& text) : x(text) {}
  void ws() {
    while (i < x.size() && (x[i] == ' ' || x[i] == '\n' || x[i] == '\r' || x[i] == '\t')) ++i;
  }
  bool eat(char c) {
    ws();
    if (i < x.size() && x[i] == c) {
      ++i;
      return true;
    }
    return false;
  }
  std::string str() {
    std::string out;
    if (!eat('"')) {
      ok = false;
      return out;
    }
    while (i < x.size() && x[i] != '"') {
      if (x[i] == '\\' && i + 1 < x.size()) {
        ++i;
        const char c = x[i];
        out += c == 'n' ? '\n' : c == 't' ? '\t' : c;
      } else {
        out += x[i];
      }
      ++i;
    }
    if (i >= x.size()) ok = false;
    ++i;
    return out;
  }
  JVal value(int depth = 0) {
    JVal v;
    ws();
    if (!ok || i >= x.size() || depth > 16) {
      ok = false;
      return v;
    }
    const char c = x[i];
    if (c == '{') {
      ++i;
      v.t = JVal::Obj;
      if (eat('}')) return v;
      do {
        std::string k = str();
        if (!ok || !eat(':')) {
          ok = false;
          return v;
        }
        v.obj.emplace_back(k, value(depth + 1));
      } while (ok && eat(','));
      if (!eat('}')) ok = false;
    } else if (c == '[') {
      ++i;
      v.t = JVal::Arr;
      if (eat(']')) return v;
      do v.arr.push_back(value(depth + 1));
      while (ok && eat(','));
      if (!eat(']')) ok = false;
    } else if (c == '"') {
      v.t = JVal::Str;
      v.s = str();
    } else if (x.compare(i, 4, "null") == 0) {
      i += 4;
    } else {
      v.t = JVal::Num;  // numbers, true, false: kept as text
      while (i < x.size() && x[i] != ',' && x[i] != '}' && x[i] != ']' && x[i] != ' ' && x[i] != '\n') v.s += x[i++];
      if (v.s.empty()) ok = false;
    }
    return v;
  }
};

static void jwrite(std::ostream& o, const JVal& v, int indent) {
  const std::string pad(indent, ' ');
  switch (v.t) {
    case JVal::Null: o << "null"; break;
    case JVal::Num: o << v.s; break;
    case JVal::Str: {
      o << '"';
      for (char c : v.s) {
        if (c == '"' || c == '\\') o << '\\';
        o << c;
      }
      o << '"';
      break;
    }
    case JVal::Arr:
      o << '[';
      for (size_t k = 0; k < v.arr.size(); ++k) {
        if (k) o << ", ";
        jwrite(o, v.arr[k], indent);
      }
      o << ']';
      break;
    case JVal::Obj:
      o << "{\n";
      for (size_t k = 0; k < v.obj.size(); ++k) {
        o << pad << "  ";
        JVal key;
        key.t = JVal::Str;
        key.s = v.obj[k].first;
        jwrite(o, key, 0);
        o << ": ";
        jwrite(o, v.obj[k].second, indent + 2);
        o << (k + 1 < v.obj.size() ? ",\n" : "\n");
      }
      o << pad << '}';
      break;
  }
}

// ---------------------------------------------------------------------------
// Targets, device key, save and load
// ---------------------------------------------------------------------------
static const TuneTarget* target_of(int t) {
  if (t == PA_TUNE_U32 || t == PA_TUNE_U64) return sort_tune_target(t);
  if (t == PA_TUNE_SEG_U32) return seg_tune_target();
  return nullptr;
}

static int device_key(std::string* out) {
  int dev = 0;
  PA_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  PA_HIP(hipGetDeviceProperties(&prop, dev));
  int rt = 0;
  PA_HIP(hipRuntimeGetVersion(&rt));
  *out = std::string(prop.gcnArchName) + "|hip " + std::to_string(rt);
  return PA_OK;
}

static std::mutex g_cache_mu;

static int read_cache(const char* path, JVal* root, bool* exists) {
  *exists = false;
  std::ifstream f(path);
  if (!f) return PA_OK;
  *exists = true;
  std::stringstream ss;
  ss << f.rdbuf();
  const std::string text = ss.str();
  JParser p(text);
  *root = p.value();
  if (!p.ok || root->t != JVal::Obj) return arg_error("tuning cache: not a JSON object");
  return PA_OK;
}

int tune_cache_save(const char* path) {
  if (!path || !*path) return arg_error("pa_tune_cache_save: empty path");
  std::string key;
  PA_TRY(device_key(&key));
  std::lock_guard<std::mutex> lk(g_cache_mu);
  JVal root;
  bool exists = false;
  if (read_cache(path, &root, &exists) != PA_OK || !exists) root = JVal();
  root.t = JVal::Obj;
  JVal fmt;
  fmt.t = JVal::Str;
  fmt.s = "pa-tune-1";
  JVal devices;
  devices.t = JVal::Obj;
  if (const JVal* d = root.get("devices"))
    if (d->t == JVal::Obj) devices = *d;
  JVal entry;
  entry.t = JVal::Obj;
  for (int t : {PA_TUNE_U32, PA_TUNE_U64, PA_TUNE_SEG_U32}) {
    const TuneTarget* T = target_of(t);
    if (!T) continue;
    JVal arr;
    arr.t = JVal::Arr;
    for (int c = 0; c < T->classes; ++c) {
      JVal v;
      const int idx = T->get(c);
      if (idx >= 0 && idx < T->count()) {
        v.t = JVal::Str;
        v.s = T->name(idx);
      }
      arr.arr.push_back(v);
    }
    entry.obj.emplace_back(T->key, arr);
  }
  bool replaced = false;
  for (auto& kv : devices.obj)
    if (kv.first == key) {
      kv.second = entry;
      replaced = true;
    }
  if (!replaced) devices.obj.emplace_back(key, entry);
  JVal out;
  out.t = JVal::Obj;
  out.obj.emplace_back("format", fmt);
  out.obj.emplace_back("devices", devices);
  const std::string tmp = std::string(path) + ".tmp";
  {
    std::ofstream f(tmp, std::ios::trunc);
    if (!f) return arg_error("pa_tune_cache_save: cannot write the cache file");
    jwrite(f, out, 0);
    f << "\n";
    if (!f) return arg_error("pa_tune_cache_save: write failed");
  }
  if (std::rename(tmp.c_str(), path) != 0) return arg_error("pa_tune_cache_save: cannot replace the cache file");
  return PA_OK;
}

int tune_cache_load(const char* path, int* applied) {
  *applied = 0;
  if (!path || !*path) return arg_error("pa_tune_cache_load: empty path");
  std::string key;
  PA_TRY(device_key(&key));
  std::lock_guard<std::mutex> lk(g_cache_mu);
  JVal root;
  bool exists = false;
  PA_TRY(read_cache(path, &root, &exists));
  if (!exists) return arg_error("pa_tune_cache_load: no such file");
  const JVal* devices = root.get("devices");
  const JVal* entry = devices ? devices->get(key) : nullptr;
  if (!entry || entry->t != JVal::Obj) return PA_OK;  // nothing for this device
  for (int t : {PA_TUNE_U32, PA_TUNE_U64, PA_TUNE_SEG_U32}) {
    const TuneTarget* T = target_of(t);
    const JVal* arr = T ? entry->get(T->key) : nullptr;
    if (!arr || arr->t != JVal::Arr) continue;
    for (int c = 0; c < T->classes && c < (int)arr->arr.size(); ++c) {
      const JVal& v = arr->arr[c];
      if (v.t != JVal::Str) continue;
      for (int i = 0; i < T->count(); ++i)
        if (T->name(i) == v.s) {
          T->set(c, i);
          ++*applied;
          break;
        }
    }
  }
  return PA_OK;
}

void tune_cache_autoload() {
  static std::once_flag once;
  std::call_once(once, [] {
    const char* path = std::getenv("PA_TUNE_CACHE");
    if (!path || !*path) return;
    std::ifstream f(path);
    if (!f) return;
    int applied = 0;
    if (tune_cache_load(path, &applied) != PA_OK) (void)hipGetLastError();
  });
}

void tune_cache_autosave() {
  const char* path = std::getenv("PA_TUNE_CACHE");
  if (path && *path && tune_cache_save(path) != PA_OK) (void)hipGetLastError();
}

}  // namespace pa

using namespace pa;

extern "C" {

int pa_sort_shape_count(int target) {
  const TuneTarget* T = target_of(target);
  return T ? T->count() : 0;
}

int pa_sort_shape_info(int target, int index, pa_sort_shape* out) {
  if (!out) return arg_error("pa_sort_shape_info: null output");
  if (!target_of(target)) return arg_error("pa_sort_shape_info: unknown target");
  return target == PA_TUNE_SEG_U32 ? seg_shape_info(index, out) : sort_shape_info(target, index, out);
}

int pa_sort_shape_choice(int target, size_t n) {
  if (!target_of(target)) return -1;
  return target == PA_TUNE_SEG_U32 ? seg_shape_choice() : sort_shape_choice(target, n);
}

int pa_sort_shape_set(int target, int size_class, int index) {
  const TuneTarget* T = target_of(target);
  if (!T) return arg_error("pa_sort_shape_set: unknown target");
  if (size_class < 0 || size_class >= T->classes) return arg_error("pa_sort_shape_set: size class out of range");
  if (index < -1 || index >= T->count()) return arg_error("pa_sort_shape_set: shape index out of range");
  T->set(size_class, index);
  return PA_OK;
}

int pa_autotune_sort(int target, size_t max_n, pa_stream_t stream) {
  if (target != PA_TUNE_U32 && target != PA_TUNE_U64) return arg_error("pa_autotune_sort: target must be u32 or u64");
  PA_TRY(autotune_sort(target, max_n, (hipStream_t)stream));
  tune_cache_autosave();
  return PA_OK;
}

int pa_tune_cache_save(const char* path) { return tune_cache_save(path); }

int pa_tune_cache_load(const char* path) {
  int applied = 0;
  return tune_cache_load(path, &applied);
}

const char* pa_tune_device_key(void) {
  static thread_local std::string key;
  key.clear();
  if (device_key(&key) != PA_OK) key.clear();
  return key.c_str();
}

}  // extern "C"
