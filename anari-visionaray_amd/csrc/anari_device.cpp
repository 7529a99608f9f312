// ANARI front end (include/anari_vrt.h) over the vrt C ABI (include/vrt.h).
//
// Object model after helium's BaseDevice as the reference uses it
// (VisionarayDevice.cpp:32-283): reference-counted objects whose parameters
// are staged by anariSetParameter and take effect at a commit; commits are
// deferred to the next anariRenderFrame (or frame query).  GPU state derived
// from committed parameters (bricked voxels, converted grids, point-location
// grids, BVHs, transfer-function tables, majorants) is built when a frame
// first needs it after a change, in dependency order, instead of in each
// object's finalize -- so commit order between objects does not matter.
// Any commit that reaches a frame's world, camera or renderer restarts its
// accumulation (Frame.cpp:499-513).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <set>
#include <string>
#include <vector>

#include "anari_vrt.h"
#include "vrt.h"

namespace {

struct Device;

bool is_object_type(ANARIDataType t) { return t >= ANARI_DEVICE && t <= ANARI_WORLD; }

size_t type_size(ANARIDataType t) {
  switch (t) {
    case ANARI_UINT8: return 1;
    case ANARI_DATA_TYPE:
    case ANARI_BOOL:
    case ANARI_INT32:
    case ANARI_UINT32:
    case ANARI_FLOAT32:
    case ANARI_UFIXED8_VEC4:
    case ANARI_UFIXED8_RGBA_SRGB: return 4;
    case ANARI_UINT32_VEC2:
    case ANARI_FLOAT32_VEC2:
    case ANARI_FLOAT32_BOX1: return 8;
    case ANARI_UINT32_VEC3:
    case ANARI_FLOAT32_VEC3: return 12;
    case ANARI_UINT32_VEC4:
    case ANARI_FLOAT32_VEC4:
    case ANARI_FLOAT32_BOX2: return 16;
    case ANARI_FLOAT32_MAT4: return 64;
    case ANARI_VOID_POINTER:
    case ANARI_FRAME_COMPLETION_CALLBACK: return sizeof(void *);
    default: return is_object_type(t) ? sizeof(void *) : 0;
  }
}

struct Object;

struct Param {
  ANARIDataType type = ANARI_UNKNOWN;
  std::vector<uint8_t> bytes;
  std::string str;
  Object *obj = nullptr;  // holds one internal reference
};

void retain_internal(Object *o);
void release_internal(Object *o);

struct Object {
  Device *dev = nullptr;
  ANARIDataType kind = ANARI_OBJECT;
  std::string subtype;
  // public (application) plus internal (parameter, commit-buffer, array
  // element) references; the object dies when the count reaches 0.  Atomic:
  // a frame completion callback may release objects on the completion thread
  // while the application thread retains and releases others.
  std::atomic<int> refs{1};
  std::map<std::string, Param> staged, committed;
  uint64_t version = 0;  // device serial of the last commit
  bool pending = false;  // in the commit buffer

  virtual ~Object() {
    for (auto *m : {&staged, &committed})
      for (auto &kv : *m)
        if (kv.second.obj) release_internal(kv.second.obj);
  }
  virtual void on_commit() {}

  const Param *param(const char *name) const {
    auto it = committed.find(name);
    return it == committed.end() ? nullptr : &it->second;
  }
  template <class T> bool get(const char *name, ANARIDataType t, T *out) const {
    const Param *p = param(name);
    if (!p || p->type != t || p->bytes.size() != sizeof(T)) return false;
    std::memcpy(out, p->bytes.data(), sizeof(T));
    return true;
  }
  float get_f(const char *name, float def) const {
    float v;
    return get(name, ANARI_FLOAT32, &v) ? v : def;
  }
  uint32_t get_u(const char *name, uint32_t def) const {
    uint32_t v;
    return get(name, ANARI_UINT32, &v) ? v : def;
  }
  int get_i(const char *name, int def) const {
    int32_t v;
    if (get(name, ANARI_INT32, &v)) return v;
    uint32_t u;
    if (get(name, ANARI_UINT32, &u)) return int(u);
    return def;
  }
  bool get_b(const char *name, bool def) const {
    int32_t v;
    return get(name, ANARI_BOOL, &v) ? v != 0 : def;
  }
  void get_v3(const char *name, float def[3]) const {
    float v[3];
    if (get(name, ANARI_FLOAT32_VEC3, &v)) std::memcpy(def, v, 12);
  }
  std::string get_s(const char *name, const char *def) const {
    const Param *p = param(name);
    return p && p->type == ANARI_STRING ? p->str : std::string(def);
  }
  Object *get_o(const char *name, ANARIDataType kind) const {
    const Param *p = param(name);
    if (!p || !p->obj) return nullptr;
    return p->obj->kind == kind ? p->obj : nullptr;
  }
};

struct Array : Object {
  ANARIDataType elem = ANARI_UNKNOWN;
  uint64_t n[3] = {1, 1, 1};
  const void *app = nullptr;
  ANARIMemoryDeleter deleter = nullptr;
  const void *user = nullptr;
  std::vector<uint8_t> owned;
  std::vector<Object *> held;  // object arrays: internal references to the elements
  bool mapped = false;

  uint64_t count() const { return n[0] * n[1] * n[2]; }
  const uint8_t *data() const {
    return app ? static_cast<const uint8_t *>(app) : owned.data();
  }
  void hold_elements() {
    for (Object *o : held) release_internal(o);
    held.clear();
    if (!is_object_type(elem)) return;
    Object *const *h = reinterpret_cast<Object *const *>(data());
    for (uint64_t i = 0; i < count(); ++i)
      if (h[i]) {
        retain_internal(h[i]);
        held.push_back(h[i]);
      }
  }
  Object *element(uint64_t i) const { return reinterpret_cast<Object *const *>(data())[i]; }
  ~Array() override {
    for (Object *o : held) release_internal(o);
    if (deleter && app) deleter(user, app);
  }
};

// --------------------------------------------------------------------------
// derived GPU state
// --------------------------------------------------------------------------
struct DevMem {
  vrt_context ctx = nullptr;
  void *p = nullptr;
  size_t bytes = 0;
  void reset() {
    if (p && ctx) vrt_free(ctx, p);
    p = nullptr;
    bytes = 0;
  }
  bool alloc(vrt_context c, size_t b) {
    reset();
    ctx = c;
    if (vrt_alloc(c, std::max<size_t>(b, 16), &p) != VRT_OK) {
      p = nullptr;
      return false;
    }
    bytes = b;
    return true;
  }
  ~DevMem() { reset(); }
};

struct Field : Object {
  uint64_t built = ~0ull;
  bool valid = false;
  vrt_volume desc{};  // field part: type, dims, bounds, cell size, data handles
  DevMem bricks;  // structuredRegular: the brick records of vrt_structured_bricks
  vrt_nanovdb nvdb = nullptr;
  vrt_unstructured tets = nullptr;
  int default_grid[3] = {16, 16, 16};
  void drop() {
    bricks.reset();
    if (nvdb) vrt_nanovdb_destroy(nvdb);
    if (tets) vrt_unstructured_destroy(tets);
    nvdb = nullptr;
    tets = nullptr;
    valid = false;
  }
  ~Field() override { drop(); }
};

struct Volume : Object {
  uint64_t built = ~0ull, field_built = ~0ull;
  bool valid = false;
  vrt_volume desc{};
  DevMem tf, ranges, majorants, mbytes, mstep;
};

struct Geometry : Object {
  uint64_t built = ~0ull;
  vrt_bvh bvh = nullptr;
  ~Geometry() override {
    if (bvh) vrt_bvh_destroy(bvh);
  }
};

// lights; an "hdri" light keeps its map and sampling tables on the device
struct Light : Object {
  uint64_t built = ~0ull;
  vrt_hdri hdri = nullptr;
  ~Light() override {
    if (hdri) vrt_hdri_destroy(hdri);
  }
};

struct Frame : Object {
  uint32_t w = 0, h = 0;
  ANARIDataType color_type = ANARI_UNKNOWN, depth_type = ANARI_UNKNOWN;
  DevMem color, accum, depth;
  // g-buffer channels (Frame.cpp:84-89, 139-162): normal, albedo (float3),
  // primitive / object / instance id (uint32); null when not requested
  DevMem gbuf[5];
  std::vector<uint8_t> host_gbuf[5];
  // temporal anti-aliasing (Frame.cpp:339-381, 515-540): motion vectors,
  // current / history colour and albedo; the previous frame's camera
  DevMem taa[5];
  bool taa_on = false;
  float taa_alpha = 0.3f;
  size_t taa_pixels = 0;
  bool has_prev_cam = false;
  vrt_camera prev_cam{};
  std::atomic<int> callbacks_pending{0};
  uint64_t built = ~0ull;
  uint64_t scene_version = ~0ull;
  int accum_id = 0, frame_id = 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  bool timed = false;
  std::vector<uint8_t> host_color;
  std::vector<float> host_depth;
  ~Frame() override {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  }
};

// channel names, ANARI element types and bytes per pixel of the g-buffer
struct GbufChannel {
  const char *name;
  ANARIDataType type;
  size_t bytes;
};
const GbufChannel kGbuf[5] = {{"channel.normal", ANARI_FLOAT32_VEC3, 12},
                              {"channel.albedo", ANARI_FLOAT32_VEC3, 12},
                              {"channel.primitiveId", ANARI_UINT32, 4},
                              {"channel.objectId", ANARI_UINT32, 4},
                              {"channel.instanceId", ANARI_UINT32, 4}};

// Frame completion callbacks run on a device-owned thread once the event
// recorded after the frame's kernels has completed.
struct Completion {
  hipEvent_t event;
  ANARIFrameCompletionCallback fn;
  const void *user;
  Frame *frame;
};

struct CompletionThread {
  // Shared by the device and the thread: a device released from inside a
  // callback is destroyed while that callback still runs on the thread, so
  // the thread keeps this block alive and leaves its loop on `orphaned`.
  struct State {
    std::mutex m;
    std::condition_variable cv;
    std::deque<Completion> q;
    bool stop = false;
    bool orphaned = false;  // the device is gone: touch nothing after the callback
    // the frame whose callback runs now, and whether that callback released
    // it (both only touched by the thread itself)
    Frame *current = nullptr;
    bool current_destroyed = false;
  };
  std::shared_ptr<State> s = std::make_shared<State>();
  std::thread t;
  void start(ANARIDevice dev) {
    std::shared_ptr<State> st = s;
    t = std::thread([st, dev] {
      for (;;) {
        Completion c;
        {
          std::unique_lock<std::mutex> lk(st->m);
          st->cv.wait(lk, [&st] { return st->stop || !st->q.empty(); });
          if (st->orphaned || st->q.empty()) return;  // stop requested and drained
          c = st->q.front();
          st->q.pop_front();
        }
        (void)hipEventSynchronize(c.event);
        (void)hipEventDestroy(c.event);
        st->current = c.frame;
        st->current_destroyed = false;
        c.fn(c.user, dev, reinterpret_cast<ANARIFrame>(c.frame));
        {
          std::lock_guard<std::mutex> lk(st->m);
          // a device destroyed by this callback outlives none of its frames,
          // so c.frame is gone too: leave without touching either
          if (st->orphaned) return;
        }
        if (!st->current_destroyed) c.frame->callbacks_pending.fetch_sub(1);
        st->current = nullptr;
      }
    });
  }
  bool on_thread() const { return t.joinable() && std::this_thread::get_id() == t.get_id(); }
  // a frame about to be destroyed by its own callback: its other queued
  // callbacks are dropped
  void forget(Frame *f) {
    if (s->current == f) s->current_destroyed = true;
    std::lock_guard<std::mutex> lk(s->m);
    for (auto it = s->q.begin(); it != s->q.end();) {
      if (it->frame == f) {
        (void)hipEventDestroy(it->event);
        it = s->q.erase(it);
      } else {
        ++it;
      }
    }
  }
  void push(const Completion &c) {
    {
      std::lock_guard<std::mutex> lk(s->m);
      s->q.push_back(c);
    }
    s->cv.notify_one();
  }
  void join() {
    if (!t.joinable()) return;
    if (on_thread()) {
      // the device released from inside a callback: the thread ends when
      // that callback returns; queued completions (none can remain: every
      // frame holds its device) are dropped
      std::lock_guard<std::mutex> lk(s->m);
      s->stop = s->orphaned = true;
      for (const Completion &c : s->q) (void)hipEventDestroy(c.event);
      s->q.clear();
      t.detach();
      return;
    }
    {
      std::lock_guard<std::mutex> lk(s->m);
      s->stop = true;
    }
    s->cv.notify_one();
    t.join();
  }
};

struct Device : Object {
  vrt_context ctx = nullptr;
  hipStream_t stream = nullptr;
  ANARIStatusCallback cb = nullptr;
  const void *cb_user = nullptr;
  std::vector<Object *> commit_buffer;
  uint64_t serial = 0;
  CompletionThread completions;
  // the instance BVH of the last world rendered, for the scene version it
  // was built at
  vrt_tls tls = nullptr;
  const Object *tls_world = nullptr;
  uint64_t tls_version = 0;

  void report(Object *src, ANARIStatusSeverity sev, ANARIStatusCode code, const char *fmt, ...)
      __attribute__((format(printf, 5, 6))) {
    if (!cb) return;
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    cb(cb_user, reinterpret_cast<ANARIDevice>(this), reinterpret_cast<ANARIObject>(src),
       src ? src->kind : ANARI_DEVICE, sev, code, buf);
  }
  void vrt_error(Object *src, const char *what) {
    report(src, ANARI_SEVERITY_ERROR, ANARI_STATUS_UNKNOWN_ERROR, "%s: %s", what,
           vrt_last_error());
  }
  void flush() {
    // commits of this batch, in the order they were requested
    std::vector<Object *> batch;
    batch.swap(commit_buffer);
    for (Object *o : batch) {
      for (auto &kv : o->committed)
        if (kv.second.obj) release_internal(kv.second.obj);
      o->committed = o->staged;
      for (auto &kv : o->committed)
        if (kv.second.obj) retain_internal(kv.second.obj);
      o->version = ++serial;
      o->pending = false;
      o->on_commit();
      release_internal(o);  // the buffer's reference
    }
  }
  ~Device() override {
    completions.join();
    if (tls) vrt_tls_destroy(tls);
    if (ctx) vrt_context_destroy(ctx);
  }
};

void retain_internal(Object *o) {
  if (o) o->refs.fetch_add(1);
}
void destroy(Object *o) {
  Device *d = o->kind == ANARI_DEVICE ? nullptr : o->dev;
  if (o->kind == ANARI_FRAME && d) {
    // frame completion callbacks name the frame: released from inside one,
    // the rest are dropped; from elsewhere, the queued ones run first
    Frame *f = static_cast<Frame *>(o);
    if (d->completions.on_thread())
      d->completions.forget(f);
    else
      while (f->callbacks_pending.load() > 0) std::this_thread::yield();
  }
  delete o;
  if (d) release_internal(d);  // every object holds its device
}
void release_internal(Object *o) {
  if (!o) return;
  if (o->refs.fetch_sub(1) == 1) destroy(o);
}

Device *as_device(ANARIDevice d) { return reinterpret_cast<Device *>(d); }
Object *as_obj(ANARIObject o) { return reinterpret_cast<Object *>(o); }

template <class T> T *make(Device *d, ANARIDataType kind, const char *subtype) {
  T *o = new T;
  o->dev = d;
  o->kind = kind;
  o->subtype = subtype ? subtype : "";
  retain_internal(d);
  return o;
}

const Array *array_param(const Object *o, const char *name) {
  const Param *p = o->param(name);
  if (!p || !p->obj) return nullptr;
  const ANARIDataType k = p->obj->kind;
  if (k != ANARI_ARRAY1D && k != ANARI_ARRAY2D && k != ANARI_ARRAY3D) return nullptr;
  return static_cast<const Array *>(p->obj);
}

// objects of an object-array parameter of element type `kind`
std::vector<Object *> objects_of(const Object *o, const char *name, ANARIDataType kind) {
  std::vector<Object *> out;
  const Array *a = array_param(o, name);
  if (!a || a->elem != kind) return out;
  for (uint64_t i = 0; i < a->count(); ++i)
    if (Object *e = a->element(i)) out.push_back(e);
  return out;
}

// --------------------------------------------------------------------------
// spatial fields (StructuredRegularField.cpp, NanoVDBField.cpp,
// UnstructuredField.cpp)
// --------------------------------------------------------------------------
bool build_field(Device *d, Field *f) {
  if (f->built == f->version) return f->valid;
  f->drop();
  f->built = f->version;
  f->desc = vrt_volume{};
  const std::string filter = f->get_s("filter", "linear");
  if (f->subtype == "structuredRegular") {
    const Array *a = array_param(f, "data");
    if (!a || a->kind != ANARI_ARRAY3D || a->elem != ANARI_FLOAT32) {
      d->report(f, ANARI_SEVERITY_WARNING, ANARI_STATUS_INVALID_ARGUMENT,
                "structuredRegular field needs a float32 Array3D 'data'");
      return false;
    }
    if (filter != "linear")
      d->report(f, ANARI_SEVERITY_WARNING, ANARI_STATUS_INVALID_ARGUMENT,
                "structuredRegular filter '%s' not supported, using linear", filter.c_str());
    float origin[3] = {0, 0, 0}, spacing[3] = {1, 1, 1};
    f->get_v3("origin", origin);
    f->get_v3("spacing", spacing);
    const int dims[3] = {int(a->n[0]), int(a->n[1]), int(a->n[2])};
    const size_t bytes = a->count() * 4;
    DevMem lin;
    const size_t brick_bytes = vrt_structured_brick_bytes(dims);
    if (brick_bytes == 0 || !lin.alloc(d->ctx, bytes) || !f->bricks.alloc(d->ctx, brick_bytes) ||
        vrt_upload(d->ctx, lin.p, a->data(), bytes) != VRT_OK ||
        vrt_structured_bricks(d->ctx, static_cast<const float *>(lin.p), dims,
                              static_cast<float *>(f->bricks.p)) != VRT_OK ||
        vrt_synchronize(d->ctx) != VRT_OK) {
      d->vrt_error(f, "structuredRegular upload");
      return false;
    }
    vrt_volume &v = f->desc;
    v.field_type = VRT_FIELD_STRUCTURED_REGULAR;
    for (int k = 0; k < 3; ++k) {
      v.dims[k] = dims[k];
      // bounds = [origin, origin + (dims-1) spacing] (StructuredRegularField.cpp:90-93)
      v.bounds_lo[k] = origin[k];
      v.bounds_hi[k] = origin[k] + (float(dims[k]) - 1.f) * spacing[k];
    }
    v.cell_size = std::min(spacing[0], std::min(spacing[1], spacing[2]));
    v.bricks = static_cast<const float *>(f->bricks.p);
    for (int k = 0; k < 3; ++k) f->default_grid[k] = 16;  // StructuredRegularField.cpp:100
  } else if (f->subtype == "nanovdb") {
    const Array *a = array_param(f, "data");
    if (!a || a->kind != ANARI_ARRAY1D) {
      d->report(f, ANARI_SEVERITY_WARNING, ANARI_STATUS_INVALID_ARGUMENT,
                "nanovdb field needs an Array1D 'data' holding the grid buffer");
      return false;
    }
    const size_t bytes = a->count() * std::max<size_t>(1, type_size(a->elem));
    // the grid needs 32-byte alignment
    std::vector<uint64_t> aligned((bytes + 7) / 8 + 4);
    uint8_t *base = reinterpret_cast<uint8_t *>(aligned.data());
    base += (32 - reinterpret_cast<uintptr_t>(base) % 32) % 32;
    std::memcpy(base, a->data(), bytes);
    if (vrt_nanovdb_create(d->ctx, base, bytes, &f->nvdb) != VRT_OK) {
      d->vrt_error(f, "nanovdb field");
      return false;
    }
    vrt_nanovdb_info info;
    vrt_nanovdb_get_info(f->nvdb, &info);
    vrt_volume &v = f->desc;
    v.field_type = VRT_FIELD_NANOVDB;
    v.nanovdb = f->nvdb;
    v.filter = filter == "nearest" ? VRT_FILTER_NEAREST : VRT_FILTER_LINEAR;
    for (int k = 0; k < 3; ++k) {
      v.bounds_lo[k] = info.bounds_lo[k];
      v.bounds_hi[k] = info.bounds_hi[k];
      f->default_grid[k] = 16;  // NanoVDBField.cpp:115
    }
    v.cell_size = 1.f;
  } else if (f->subtype == "unstructured") {
    const Array *pos = array_param(f, "vertex.position"), *vd = array_param(f, "vertex.data");
    const Array *idx = array_param(f, "index"), *ci = array_param(f, "cell.index");
    const Array *ct = array_param(f, "cell.type"), *cd = array_param(f, "cell.data");
    if (!pos || pos->elem != ANARI_FLOAT32_VEC3 || !idx || idx->elem != ANARI_UINT32 || !ci ||
        ci->elem != ANARI_UINT32 || !ct || ct->elem != ANARI_UINT8 ||
        (vd && vd->elem != ANARI_FLOAT32) || (cd && cd->elem != ANARI_FLOAT32)) {
      d->report(f, ANARI_SEVERITY_WARNING, ANARI_STATUS_INVALID_ARGUMENT,
                "unstructured field needs float3 'vertex.position', uint32 'index', "
                "'cell.index', uint8 'cell.type' and float 'vertex.data' or 'cell.data'");
      return false;
    }
    if (ct->count() != ci->count()) {
      d->report(f, ANARI_SEVERITY_WARNING, ANARI_STATUS_INVALID_ARGUMENT,
                "'cell.type' and 'cell.index' differ in length");
      return false;
    }
    if (vrt_unstructured_create(
            d->ctx, reinterpret_cast<const float *>(pos->data()),
            vd ? reinterpret_cast<const float *>(vd->data()) : nullptr, int64_t(pos->count()),
            reinterpret_cast<const uint32_t *>(idx->data()), int64_t(idx->count()),
            reinterpret_cast<const uint32_t *>(ci->data()), ct->data(),
            cd ? reinterpret_cast<const float *>(cd->data()) : nullptr, int64_t(ci->count()),
            &f->tets) != VRT_OK) {
      d->vrt_error(f, "unstructured field");
      return false;
    }
    vrt_unstructured_info info;
    vrt_unstructured_get_info(f->tets, &info);
    vrt_volume &v = f->desc;
    v.field_type = VRT_FIELD_UNSTRUCTURED;
    v.unstructured = f->tets;
    for (int k = 0; k < 3; ++k) {
      v.bounds_lo[k] = info.bounds_lo[k];
      v.bounds_hi[k] = info.bounds_hi[k];
      f->default_grid[k] = 64;  // UnstructuredField.cpp:380
    }
    v.cell_size = info.cell_size;
  } else {
    d->report(f, ANARI_SEVERITY_WARNING, ANARI_STATUS_INVALID_ARGUMENT,
              "unknown spatial field subtype '%s'", f->subtype.c_str());
    return false;
  }
  f->valid = true;
  return true;
}

// --------------------------------------------------------------------------
// transferFunction1D volume (TransferFunction1D.cpp:23-130)
// --------------------------------------------------------------------------
bool build_volume(Device *d, Volume *v) {
  Field *f = static_cast<Field *>(v->get_o("value", ANARI_SPATIAL_FIELD));
  if (!f) f = static_cast<Field *>(v->get_o("field", ANARI_SPATIAL_FIELD));
  if (v->subtype != "transferFunction1D") {
    if (v->built != v->version)
      d->report(v, ANARI_SEVERITY_WARNING, ANARI_STATUS_INVALID_ARGUMENT,
                "unknown volume subtype '%s'", v->subtype.c_str());
    v->built = v->version;
    return v->valid = false;
  }
  if (!f) {
    if (v->built != v->version)
      d->report(v, ANARI_SEVERITY_WARNING, ANARI_STATUS_INVALID_ARGUMENT,
                "no spatial field provided to transferFunction1D volume");
    v->built = v->version;
    return v->valid = false;
  }
  if (!build_field(d, f)) return v->valid = false;
  if (v->built == v->version && v->field_built == f->built) return v->valid;
  v->built = v->version;
  v->field_built = f->built;
  v->valid = false;

  // table: color (array of float3/float4, or uniform), opacity (array or
  // uniform), uniform color.w times uniform opacity (:38-47)
  float uni[4] = {1, 1, 1, 1};
  float c4[4];
  if (v->get("color", ANARI_FLOAT32_VEC4, &c4)) std::memcpy(uni, c4, 16);
  else v->get_v3("color", uni);
  uni[3] = v->get_f("opacity", 1.f) * uni[3];
  const Array *ca = array_param(v, "color"), *oa = array_param(v, "opacity");
  if (ca && ca->elem != ANARI_FLOAT32_VEC3 && ca->elem != ANARI_FLOAT32_VEC4) ca = nullptr;
  if (oa && oa->elem != ANARI_FLOAT32) oa = nullptr;
  const float *colors = ca ? reinterpret_cast<const float *>(ca->data()) : nullptr;
  const int nch = ca && ca->elem == ANARI_FLOAT32_VEC3 ? 3 : 4;
  const float *opac = oa ? reinterpret_cast<const float *>(oa->data()) : nullptr;
  const int n = vrt_tf_build(colors, ca ? int(ca->count()) : 0, nch, opac,
                             oa ? int(oa->count()) : 0, uni, nullptr);
  if (n < 1) {
    d->vrt_error(v, "transfer function");
    return false;
  }
  std::vector<float> table(size_t(n) * 4);
  vrt_tf_build(colors, ca ? int(ca->count()) : 0, nch, opac, oa ? int(oa->count()) : 0, uni,
               table.data());
  float range[2] = {0.f, 1.f};
  if (!v->get("valueRange", ANARI_FLOAT32_BOX1, &range)) v->get("valueRange", ANARI_FLOAT32_VEC2, &range);
  float ud = v->get_f("unitDistance", 1.f);
  if (!(ud > 0.f)) {
    d->report(v, ANARI_SEVERITY_WARNING, ANARI_STATUS_INVALID_ARGUMENT,
              "invalid/zero unitDistance on transferFunction1D volume, setting to 1.0");
    ud = 1.f;
  }
  int grid[3] = {f->default_grid[0], f->default_grid[1], f->default_grid[2]};
  uint32_t g3[3];
  if (v->get("majorantGrid", ANARI_UINT32_VEC3, &g3))  // extension
    for (int k = 0; k < 3; ++k) grid[k] = int(std::max(1u, std::min(g3[k], 1024u)));
  const size_t nmc = size_t(grid[0]) * grid[1] * grid[2];

  vrt_volume &vd = v->desc;
  vd = f->desc;
  if (!v->tf.alloc(d->ctx, table.size() * 4) ||
      vrt_upload(d->ctx, v->tf.p, table.data(), table.size() * 4) != VRT_OK) {
    d->vrt_error(v, "transfer function upload");
    return false;
  }
  vd.tf_rgba = static_cast<const float *>(v->tf.p);
  vd.tf_size = n;
  vd.value_lo = range[0];
  vd.value_hi = range[1];
  vd.unit_distance = ud;
  for (int k = 0; k < 3; ++k) vd.grid_dims[k] = grid[k];
  vd.affine_inv[0] = vd.affine_inv[4] = vd.affine_inv[8] = 1.f;
  if (!v->ranges.alloc(d->ctx, nmc * 8) || !v->majorants.alloc(d->ctx, nmc * 4) ||
      !v->mbytes.alloc(d->ctx, nmc) || !v->mstep.alloc(d->ctx, 4)) {
    d->vrt_error(v, "majorant grid allocation");
    return false;
  }
  vd.majorants = static_cast<const float *>(v->majorants.p);
  vd.majorant_bytes = static_cast<const uint8_t *>(v->mbytes.p);
  vd.majorant_step = static_cast<const float *>(v->mstep.p);
  if (vrt_build_value_ranges(d->ctx, &vd, static_cast<float *>(v->ranges.p)) != VRT_OK ||
      vrt_compute_majorants(d->ctx, &vd, static_cast<const float *>(v->ranges.p),
                            static_cast<float *>(v->majorants.p),
                            static_cast<uint8_t *>(v->mbytes.p),
                            static_cast<float *>(v->mstep.p)) != VRT_OK ||
      vrt_synchronize(d->ctx) != VRT_OK) {
    d->vrt_error(v, "majorant grid");
    return false;
  }
  return v->valid = true;
}

// --------------------------------------------------------------------------
// triangle geometry (Triangle.cpp) -> BVH
// --------------------------------------------------------------------------
vrt_bvh build_geometry(Device *d, Geometry *g) {
  if (g->built == g->version) return g->bvh;
  g->built = g->version;
  if (g->bvh) vrt_bvh_destroy(g->bvh);
  g->bvh = nullptr;
  if (g->subtype != "triangle") {
    d->report(g, ANARI_SEVERITY_WARNING, ANARI_STATUS_INVALID_ARGUMENT,
              "geometry subtype '%s' not supported", g->subtype.c_str());
    return nullptr;
  }
  const Array *pos = array_param(g, "vertex.position"), *idx = array_param(g, "primitive.index");
  if (!pos || pos->elem != ANARI_FLOAT32_VEC3 || (idx && idx->elem != ANARI_UINT32_VEC3)) {
    d->report(g, ANARI_SEVERITY_WARNING, ANARI_STATUS_INVALID_ARGUMENT,
              "triangle geometry needs float3 'vertex.position' (and uint3 'primitive.index')");
    return nullptr;
  }
  const int64_t ntri = idx ? int64_t(idx->count()) : int64_t(pos->count() / 3);
  // extension "bvhBuilder": "sah" (default: the reference's host binned-SAH
  // build, Triangle.cpp:200-202) or "lbvh" (built on the GPU from uploaded
  // copies of the arrays, for geometry re-committed every frame)
  const std::string builder = g->get_s("bvhBuilder", "sah");
  int rc = VRT_OK;
  if (builder == "lbvh") {
    const size_t pbytes = pos->count() * 12, ibytes = idx ? idx->count() * 12 : 0;
    void *dp = nullptr, *di = nullptr;
    hipError_t e = hipMalloc(&dp, pbytes);
    if (e == hipSuccess && idx) e = hipMalloc(&di, ibytes);
    if (e == hipSuccess) e = hipMemcpyAsync(dp, pos->data(), pbytes, hipMemcpyHostToDevice, d->stream);
    if (e == hipSuccess && idx)
      e = hipMemcpyAsync(di, idx->data(), ibytes, hipMemcpyHostToDevice, d->stream);
    if (e == hipSuccess) {
      rc = vrt_bvh_build_gpu(d->ctx, static_cast<const float *>(dp), int64_t(pos->count()),
                             static_cast<const uint32_t *>(di), ntri, &g->bvh);
    } else {
      d->report(g, ANARI_SEVERITY_ERROR, ANARI_STATUS_OUT_OF_MEMORY, "BVH upload: %s",
                hipGetErrorString(e));
      g->bvh = nullptr;
    }
    (void)hipStreamSynchronize(d->stream);
    (void)hipFree(dp);
    (void)hipFree(di);
    if (e != hipSuccess) return nullptr;
  } else {
    if (builder != "sah")
      d->report(g, ANARI_SEVERITY_WARNING, ANARI_STATUS_INVALID_ARGUMENT,
                "unknown bvhBuilder '%s': using 'sah'", builder.c_str());
    rc = vrt_bvh_build(d->ctx, reinterpret_cast<const float *>(pos->data()),
                       int64_t(pos->count()),
                       idx ? reinterpret_cast<const uint32_t *>(idx->data()) : nullptr, ntri,
                       &g->bvh);
  }
  if (rc != VRT_OK) {
    d->vrt_error(g, "triangle BVH");
    g->bvh = nullptr;
  }
  return g->bvh;
}

// --------------------------------------------------------------------------
// scene assembly: world -> instances -> groups (World.cpp, Instance.cpp,
// Group.cpp); transforms are ANARI column-major mat4
// --------------------------------------------------------------------------
struct Affine {
  float m[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};  // row-major 3x4 [L | t]
  bool identity() const {
    static const float id[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
    return std::memcmp(m, id, sizeof(m)) == 0;
  }
  void point(const float p[3], float o[3]) const {
    for (int r = 0; r < 3; ++r) o[r] = m[4 * r] * p[0] + m[4 * r + 1] * p[1] + m[4 * r + 2] * p[2] + m[4 * r + 3];
  }
  void vector(const float p[3], float o[3]) const {
    for (int r = 0; r < 3; ++r) o[r] = m[4 * r] * p[0] + m[4 * r + 1] * p[1] + m[4 * r + 2] * p[2];
  }
};

struct SceneBuild {
  vrt_scene scene{};
  std::vector<vrt_volume> volumes;
  std::vector<vrt_light> lights;
  std::vector<vrt_surface> surfaces;
  std::vector<vrt_instance> instances;  // surfaces under transforms, and beyond the surface slots
  uint64_t version = 0;  // newest commit among the objects used
};

// inst_id: the instance's "id" (the world's for its own objects, World.cpp:96)
void add_group(Device *d, const Object *grp, const Affine &x, uint32_t inst_id, SceneBuild &sb) {
  sb.version = std::max(sb.version, grp->version);
  for (Object *o : objects_of(grp, "volume", ANARI_VOLUME)) {
    Volume *v = static_cast<Volume *>(o);
    sb.version = std::max(sb.version, v->version);
    if (Object *f = v->get_o("value", ANARI_SPATIAL_FIELD)) sb.version = std::max(sb.version, f->version);
    if (!build_volume(d, v)) continue;
    if (sb.volumes.size() >= VRT_MAX_VOLUMES) {
      d->report(v, ANARI_SEVERITY_ERROR, ANARI_STATUS_INVALID_OPERATION,
                "more than %d volumes in a world: this volume is not rendered",
                VRT_MAX_VOLUMES);
      continue;
    }
    vrt_volume vd = v->desc;
    // affineInv / transInv of the instance (DeviceCopyableObjects.h:1659-1680)
    double L[9];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) L[3 * r + c] = x.m[4 * r + c];
    const double det = L[0] * (L[4] * L[8] - L[5] * L[7]) - L[1] * (L[3] * L[8] - L[5] * L[6]) +
                       L[2] * (L[3] * L[7] - L[4] * L[6]);
    if (det == 0.0) {
      d->report(v, ANARI_SEVERITY_WARNING, ANARI_STATUS_INVALID_ARGUMENT,
                "singular instance transform: volume ignored");
      continue;
    }
    const double inv[9] = {(L[4] * L[8] - L[5] * L[7]) / det, (L[2] * L[7] - L[1] * L[8]) / det,
                           (L[1] * L[5] - L[2] * L[4]) / det, (L[5] * L[6] - L[3] * L[8]) / det,
                           (L[0] * L[8] - L[2] * L[6]) / det, (L[2] * L[3] - L[0] * L[5]) / det,
                           (L[3] * L[7] - L[4] * L[6]) / det, (L[1] * L[6] - L[0] * L[7]) / det,
                           (L[0] * L[4] - L[1] * L[3]) / det};
    for (int i = 0; i < 9; ++i) vd.affine_inv[i] = float(inv[i]);
    for (int k = 0; k < 3; ++k) vd.trans_inv[k] = -x.m[4 * k + 3];
    vd.object_id = v->get_u("id", ~0u);  // Volume.cpp:25
    vd.instance_id = inst_id;
    sb.volumes.push_back(vd);
  }
  for (Object *o : objects_of(grp, "surface", ANARI_SURFACE)) {
    sb.version = std::max(sb.version, o->version);
    Geometry *g = static_cast<Geometry *>(o->get_o("geometry", ANARI_GEOMETRY));
    Object *m = o->get_o("material", ANARI_MATERIAL);
    if (!g) continue;
    sb.version = std::max(sb.version, g->version);
    if (m) sb.version = std::max(sb.version, m->version);
    vrt_bvh bvh = build_geometry(d, g);
    if (!bvh) continue;
    float c[4] = {1, 1, 1, 1};  // Matte.h:17
    if (m) {
      if (m->subtype != "matte")
        d->report(m, ANARI_SEVERITY_WARNING, ANARI_STATUS_INVALID_OPERATION,
                  "material '%s' not supported: rendered as matte", m->subtype.c_str());
      if (!m->get("color", ANARI_FLOAT32_VEC4, &c)) m->get_v3("color", c);
    }
    if (x.identity() && sb.surfaces.size() < VRT_MAX_SURFACES) {
      vrt_surface s{};
      s.bvh = bvh;
      for (int k = 0; k < 3; ++k) s.color[k] = c[k];
      s.object_id = o->get_u("id", ~0u);  // Surface.cpp:21
      s.instance_id = inst_id;
      sb.surfaces.push_back(s);
    } else {
      // a surface under a transform (or beyond the surface slots): an
      // instance of the world's instance BVH (VisionarayScene.cpp:69-117)
      vrt_instance in{};
      in.bvh = bvh;
      std::memcpy(in.transform, x.m, sizeof(in.transform));
      for (int k = 0; k < 3; ++k) in.color[k] = c[k];
      in.object_id = o->get_u("id", ~0u);
      in.instance_id = inst_id;
      sb.instances.push_back(in);
    }
  }
  for (Object *o : objects_of(grp, "light", ANARI_LIGHT)) {
    sb.version = std::max(sb.version, o->version);
    if (sb.lights.size() >= VRT_MAX_LIGHTS) {
      d->report(o, ANARI_SEVERITY_ERROR, ANARI_STATUS_INVALID_OPERATION,
                "more than %d lights in a world: this light is not rendered", VRT_MAX_LIGHTS);
      continue;
    }
    vrt_light L{};
    L.visible = o->get_b("visible", true) ? 1 : 0;
    float color[3] = {1, 1, 1};
    o->get_v3("color", color);
    std::memcpy(L.color, color, 12);
    if (o->subtype == "directional") {
      // Directional.cpp:20-26 (the kernels' scene set-up normalizes it)
      float dir[3] = {0, 0, -1};
      o->get_v3("direction", dir);
      x.vector(dir, L.direction);
      L.type = VRT_LIGHT_DIRECTIONAL;
      L.intensity = std::max(0.f, o->get_f("irradiance", 1.f));
    } else if (o->subtype == "quad") {
      // Quad.cpp:19-45
      float p[3] = {0, 0, 0}, e1[3] = {1, 0, 0}, e2[3] = {0, 1, 0};
      o->get_v3("position", p);
      o->get_v3("edge1", e1);
      o->get_v3("edge2", e2);
      x.point(p, L.position);
      x.vector(e1, L.edge1);
      x.vector(e2, L.edge2);
      L.type = VRT_LIGHT_QUAD;
      L.intensity = std::max(0.f, o->get_f("intensity", 1.f));
    } else if (o->subtype == "hdri") {
      // HDRI.cpp:102-162: "radiance" (2D float3 map), "up", "direction",
      // "scale"; not moved by instance transforms (xfmLight)
      Light *hl = static_cast<Light *>(o);
      const Array *rad = array_param(o, "radiance");
      if (!rad || rad->kind != ANARI_ARRAY2D || rad->elem != ANARI_FLOAT32_VEC3) {
        d->report(o, ANARI_SEVERITY_WARNING, ANARI_STATUS_INVALID_ARGUMENT,
                  "missing required parameter 'radiance' on 'hdri' light");
        continue;
      }
      sb.version = std::max(sb.version, rad->version);
      const uint64_t ver = std::max(o->version, rad->version);
      if (hl->built != ver) {
        hl->built = ver;
        if (hl->hdri) vrt_hdri_destroy(hl->hdri);
        hl->hdri = nullptr;
        if (vrt_hdri_create(d->ctx, reinterpret_cast<const float *>(rad->data()), int(rad->n[0]),
                            int(rad->n[1]), &hl->hdri) != VRT_OK) {
          d->vrt_error(o, "hdri light");
          hl->hdri = nullptr;
        }
      }
      if (!hl->hdri) continue;
      L.type = VRT_LIGHT_HDRI;
      L.hdri = hl->hdri;
      float up[3] = {1, 0, 0}, dir[3] = {0, 0, 1};
      o->get_v3("up", up);
      o->get_v3("direction", dir);
      std::memcpy(L.up, up, 12);
      std::memcpy(L.direction, dir, 12);
      L.scale = o->get_f("scale", 1.f);
    } else {
      d->report(o, ANARI_SEVERITY_WARNING, ANARI_STATUS_INVALID_OPERATION,
                "light subtype '%s' not supported: ignored", o->subtype.c_str());
      continue;
    }
    sb.lights.push_back(L);
  }
}

bool assemble(Device *d, const Object *world, SceneBuild &sb) {
  sb.version = world->version;
  // the world's own surfaces/volumes/lights form a group under identity
  // (World.cpp:56-97)
  add_group(d, world, Affine{}, world->get_u("id", ~0u), sb);
  for (Object *inst : objects_of(world, "instance", ANARI_INSTANCE)) {
    sb.version = std::max(sb.version, inst->version);
    const Object *g = inst->get_o("group", ANARI_GROUP);
    if (!g) continue;
    Affine x;
    float m4[16];
    if (inst->get("transform", ANARI_FLOAT32_MAT4, &m4))
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 4; ++c) x.m[4 * r + c] = m4[4 * c + r];  // column-major
    add_group(d, g, x, inst->get_u("id", ~0u), sb);
  }
  sb.scene.num_volumes = int(sb.volumes.size());
  std::copy(sb.volumes.begin(), sb.volumes.end(), sb.scene.volumes);
  sb.scene.num_lights = int(sb.lights.size());
  std::copy(sb.lights.begin(), sb.lights.end(), sb.scene.lights);
  sb.scene.num_surfaces = int(sb.surfaces.size());
  std::copy(sb.surfaces.begin(), sb.surfaces.end(), sb.scene.surfaces);
  sb.scene.tls = nullptr;
  if (!sb.instances.empty()) {
    if (!d->tls || d->tls_world != world || d->tls_version != sb.version) {
      if (d->tls) vrt_tls_destroy(d->tls);
      d->tls = nullptr;
      d->tls_world = world;
      d->tls_version = sb.version;
      if (vrt_tls_build(d->ctx, sb.instances.data(), int(sb.instances.size()), &d->tls) !=
          VRT_OK) {
        d->vrt_error(const_cast<Object *>(world), "instance BVH");
        d->tls = nullptr;
      }
    }
    sb.scene.tls = d->tls;
  }
  return true;
}

// --------------------------------------------------------------------------
// frame (Frame.cpp:215-397)
// --------------------------------------------------------------------------
bool build_frame(Device *d, Frame *fr) {
  if (fr->built == fr->version) return true;
  fr->built = fr->version;
  uint32_t size[2] = {10, 10};
  fr->get("size", ANARI_UINT32_VEC2, &size);
  int32_t ct = ANARI_UNKNOWN, dt = ANARI_UNKNOWN;
  fr->get("channel.color", ANARI_DATA_TYPE, &ct);
  fr->get("channel.depth", ANARI_DATA_TYPE, &dt);
  if (ct != ANARI_UNKNOWN && ct != ANARI_FLOAT32_VEC4 && ct != ANARI_UFIXED8_VEC4 &&
      ct != ANARI_UFIXED8_RGBA_SRGB) {
    d->report(fr, ANARI_SEVERITY_WARNING, ANARI_STATUS_INVALID_ARGUMENT,
              "unsupported channel.color type %d, using FLOAT32_VEC4", ct);
    ct = ANARI_FLOAT32_VEC4;
  }
  if (size[0] < 1 || size[1] < 1 || size[0] > 16384 || size[1] > 16384) {
    d->report(fr, ANARI_SEVERITY_ERROR, ANARI_STATUS_INVALID_ARGUMENT, "frame size %ux%u",
              size[0], size[1]);
    return false;
  }
  fr->w = size[0];
  fr->h = size[1];
  fr->color_type = ct;
  fr->depth_type = dt == ANARI_FLOAT32 ? ANARI_FLOAT32 : ANARI_UNKNOWN;
  const size_t np = size_t(fr->w) * fr->h;
  const size_t cbytes = np * (ct == ANARI_FLOAT32_VEC4 || ct == ANARI_UNKNOWN ? 16 : 4);
  if (!fr->color.alloc(d->ctx, cbytes) || !fr->accum.alloc(d->ctx, np * 16) ||
      (fr->depth_type == ANARI_FLOAT32 && !fr->depth.alloc(d->ctx, np * 4))) {
    d->vrt_error(fr, "frame allocation");
    return false;
  }
  if (fr->depth_type != ANARI_FLOAT32) fr->depth.reset();
  for (int k = 0; k < 5; ++k) {
    int32_t t = ANARI_UNKNOWN;
    fr->get(kGbuf[k].name, ANARI_DATA_TYPE, &t);
    if (t == ANARI_UNKNOWN) {
      fr->gbuf[k].reset();
      continue;
    }
    if (t != kGbuf[k].type) {
      d->report(fr, ANARI_SEVERITY_WARNING, ANARI_STATUS_INVALID_ARGUMENT,
                "%s must be of type %d: channel ignored", kGbuf[k].name, kGbuf[k].type);
      fr->gbuf[k].reset();
      continue;
    }
    if (!fr->gbuf[k].alloc(d->ctx, np * kGbuf[k].bytes) ||
        vrt_memset(d->ctx, fr->gbuf[k].p, 0, np * kGbuf[k].bytes) != VRT_OK) {
      d->vrt_error(fr, "frame allocation");
      return false;
    }
  }
  fr->scene_version = ~0ull;  // a new size restarts accumulation
  return true;
}

void render(Device *d, Frame *fr) {
  d->flush();
  if (!build_frame(d, fr)) return;
  Object *world = fr->get_o("world", ANARI_WORLD);
  Object *cam = fr->get_o("camera", ANARI_CAMERA);
  Object *ren = fr->get_o("renderer", ANARI_RENDERER);
  if (!world || !cam || !ren) {
    d->report(fr, ANARI_SEVERITY_ERROR, ANARI_STATUS_INVALID_OPERATION,
              "frame needs 'world', 'camera' and 'renderer'");
    return;
  }
  SceneBuild sb;
  assemble(d, world, sb);
  const uint64_t ver = std::max({sb.version, cam->version, ren->version, fr->version});

  // camera: perspective (camera/Camera.cpp:32-40, Perspective.cpp:14-44)
  vrt_camera c{};
  float pos[3] = {0, 0, 0}, dir[3] = {0, 0, 1}, up[3] = {0, 1, 0};
  cam->get_v3("position", pos);
  cam->get_v3("direction", dir);
  cam->get_v3("up", up);
  std::memcpy(c.position, pos, 12);
  std::memcpy(c.direction, dir, 12);
  std::memcpy(c.up, up, 12);
  c.fovy = cam->get_f("fovy", 60.f * 3.14159265358979f / 180.f);
  c.aspect = cam->get_f("aspect", 1.f);
  float region[4] = {0, 0, 1, 1}, shutter[2] = {0.5f, 0.5f};
  cam->get("imageRegion", ANARI_FLOAT32_BOX2, &region);
  cam->get("shutter", ANARI_FLOAT32_BOX1, &shutter);
  std::memcpy(c.image_region, region, 16);
  std::memcpy(c.shutter, shutter, 8);
  c.aperture_radius = cam->get_f("apertureRadius", 0.f);
  c.focus_distance = cam->get_f("focusDistance", 1.f);
  if (cam->subtype != "perspective")
    d->report(cam, ANARI_SEVERITY_WARNING, ANARI_STATUS_INVALID_OPERATION,
              "camera '%s' not supported: rendered as perspective", cam->subtype.c_str());

  // renderer (Renderer.cpp:42-53, Pathtrace.cpp:17-26, Raycast.cpp:17-20)
  vrt_params p{};
  float bg[4] = {0, 0, 0, 1};
  ren->get("background", ANARI_FLOAT32_VEC4, &bg);
  std::memcpy(p.background, bg, 16);
  float ac[3] = {1, 1, 1};
  ren->get_v3("ambientColor", ac);
  std::memcpy(p.ambient_color, ac, 12);
  p.ambient_radiance = ren->get_f("ambientRadiance", 0.2f);
  p.occlusion_distance = ren->get_f("ambientOcclusionDistance", 1e20f);
  p.ambient_samples = std::min(std::max(ren->get_i("ambientSamples", 1), 0), 256);
  p.max_bounce = std::min(std::max(ren->get_i("maxBounce", 7), 0), 256);
  p.pixel_samples = std::min(std::max(ren->get_i("pixelSamples", 1), 1), 256);
  p.volume_sampling_rate = ren->get_f("volumeSamplingRate", 0.5f);
  p.phase_g = ren->get_f("phaseG", 0.f);  // extension
  p.gradient_shading = ren->get_b("gradientShading", false) ? 1 : 0;  // Renderer.cpp:50
  // extension: shadow / AO transmittance by ratio tracking
  p.ratio_tracking = ren->get_b("ratioTracking", false) ? 1 : 0;
  if (const Array *cp = array_param(ren, "clipPlane")) {
    if (cp->elem == ANARI_FLOAT32_VEC4) {
      p.num_clip_planes = int(std::min<uint64_t>(cp->count(), VRT_MAX_CLIP_PLANES));
      std::memcpy(p.clip_planes, cp->data(), size_t(p.num_clip_planes) * 16);
    }
  }
  const bool pathtrace = ren->subtype == "pathtrace";
  // path tracer "taa" / "taaAlpha" (Pathtrace.cpp:24-25)
  const bool taa = pathtrace && ren->get_b("taa", false);
  const float taa_alpha = ren->get_f("taaAlpha", 0.3f);
  if (!pathtrace && ren->subtype != "raycast" && ren->subtype != "default")
    d->report(ren, ANARI_SEVERITY_WARNING, ANARI_STATUS_INVALID_OPERATION,
              "renderer '%s' not supported: using raycast", ren->subtype.c_str());

  vrt_frame f{};
  f.width = int(fr->w);
  f.height = int(fr->h);
  f.color_type = fr->color_type == ANARI_UNKNOWN ? VRT_TYPE_FLOAT32_VEC4 : fr->color_type;
  f.color = fr->color.p;
  f.accum = static_cast<float *>(fr->accum.p);
  f.depth = fr->depth_type == ANARI_FLOAT32 ? static_cast<float *>(fr->depth.p) : nullptr;
  if (f.depth) {  // written only together with depth (Frame::fillGBuffer)
    f.normal = static_cast<float *>(fr->gbuf[0].p);
    f.albedo = static_cast<float *>(fr->gbuf[1].p);
    f.primitive_id = static_cast<uint32_t *>(fr->gbuf[2].p);
    f.object_id = static_cast<uint32_t *>(fr->gbuf[3].p);
    f.instance_id = static_cast<uint32_t *>(fr->gbuf[4].p);
  }
  if (taa) {
    // checkTAAReset (Frame.cpp:515-540): fresh zeroed buffers when TAA is
    // switched on, the size changes or alpha changes
    const size_t np = size_t(fr->w) * fr->h;
    if (!fr->taa_on || fr->taa_pixels != np || fr->taa_alpha != taa_alpha) {
      const size_t bytes[5] = {np * 16, np * 16, np * 16, np * 12, np * 12};
      for (int k = 0; k < 5; ++k)
        if (!fr->taa[k].alloc(d->ctx, bytes[k]) ||
            vrt_memset(d->ctx, fr->taa[k].p, 0, bytes[k]) != VRT_OK) {
          d->vrt_error(fr, "taa buffers");
          return;
        }
      fr->taa_pixels = np;
      fr->taa_alpha = taa_alpha;
    }
    fr->taa_on = true;
    f.motion_vec = static_cast<float *>(fr->taa[0].p);
    f.taa_curr = static_cast<float *>(fr->taa[1].p);
    f.taa_prev = static_cast<float *>(fr->taa[2].p);
    f.taa_curr_albedo = static_cast<float *>(fr->taa[3].p);
    f.taa_prev_albedo = static_cast<float *>(fr->taa[4].p);
    p.taa = 1;
    p.taa_alpha = taa_alpha;
    p.has_prev_camera = fr->has_prev_cam ? 1 : 0;
    p.prev_camera = fr->prev_cam;
  } else {
    fr->taa_on = false;
  }
  if (ver != fr->scene_version) {  // Frame.cpp:499-513
    fr->scene_version = ver;
    fr->accum_id = 0;
  }
  if (fr->accum_id == 0) {
    if (vrt_frame_reset(d->ctx, &f) != VRT_OK) {
      d->vrt_error(fr, "frame reset");
      return;
    }
  }
  if (!fr->e0) {
    (void)hipEventCreate(&fr->e0);
    (void)hipEventCreate(&fr->e1);
  }
  (void)hipEventRecord(fr->e0, d->stream);
  const int rc = pathtrace ? vrt_pathtrace(d->ctx, &sb.scene, &c, &p, &f, fr->accum_id, fr->frame_id)
                           : vrt_raycast(d->ctx, &sb.scene, &c, &p, &f, fr->frame_id);
  (void)hipEventRecord(fr->e1, d->stream);
  fr->timed = true;
  if (rc != VRT_OK) {
    d->vrt_error(fr, pathtrace ? "pathtrace" : "raycast");
    return;
  }
  ++fr->accum_id;
  ++fr->frame_id;
  // the matrices of this frame become the previous ones (Frame.cpp:333-334)
  fr->prev_cam = c;
  fr->has_prev_cam = true;
  // frameCompletionCallback (Frame.cpp:90-93, 383-384)
  ANARIFrameCompletionCallback fn = nullptr;
  const void *user = nullptr;
  fr->get("frameCompletionCallback", ANARI_FRAME_COMPLETION_CALLBACK, &fn);
  fr->get("frameCompletionCallbackUserData", ANARI_VOID_POINTER, &user);
  if (fn) {
    hipEvent_t ev = nullptr;
    if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess ||
        hipEventRecord(ev, d->stream) != hipSuccess) {
      d->report(fr, ANARI_SEVERITY_ERROR, ANARI_STATUS_UNKNOWN_ERROR,
                "frame completion event failed");
      if (ev) (void)hipEventDestroy(ev);
      return;
    }
    if (!d->completions.t.joinable()) d->completions.start(reinterpret_cast<ANARIDevice>(d));
    fr->callbacks_pending.fetch_add(1);
    d->completions.push({ev, fn, user, fr});
  }
}

}  // namespace

// ==========================================================================
// C API
// ==========================================================================
extern "C" {

ANARIDevice anariNewVisionarayDevice(ANARIStatusCallback cb, const void *user) {
  Device *d = new Device;
  d->dev = d;
  d->kind = ANARI_DEVICE;
  d->subtype = "visionaray_amd";
  d->cb = cb;
  d->cb_user = user;
  void *stream = nullptr;
  if (vrt_context_create(0, &d->ctx) != VRT_OK ||
      vrt_context_stream(d->ctx, &stream) != VRT_OK) {
    d->report(nullptr, ANARI_SEVERITY_FATAL_ERROR, ANARI_STATUS_UNSUPPORTED_DEVICE,
              "no usable GPU: %s", vrt_last_error());
    delete d;
    return nullptr;
  }
  d->stream = static_cast<hipStream_t>(stream);
  return reinterpret_cast<ANARIDevice>(d);
}

static Object *new_array(ANARIDevice dh, const void *app, ANARIMemoryDeleter del,
                         const void *user, ANARIDataType t, uint64_t n1, uint64_t n2,
                         uint64_t n3, ANARIDataType kind) {
  Device *d = as_device(dh);
  if (!d) return nullptr;
  const size_t es = type_size(t);
  if (es == 0 || n1 == 0 || n2 == 0 || n3 == 0) {
    d->report(nullptr, ANARI_SEVERITY_ERROR, ANARI_STATUS_INVALID_ARGUMENT,
              "array of element type %d and size %llu x %llu x %llu not supported", t,
              (unsigned long long)n1, (unsigned long long)n2, (unsigned long long)n3);
    return nullptr;
  }
  Array *a = make<Array>(d, kind, "");
  a->elem = t;
  a->n[0] = n1;
  a->n[1] = n2;
  a->n[2] = n3;
  a->app = app;
  a->deleter = del;
  a->user = user;
  if (!app) a->owned.assign(size_t(a->count()) * es, 0);
  a->hold_elements();
  a->version = ++d->serial;
  return a;
}

ANARIArray1D anariNewArray1D(ANARIDevice d, const void *app, ANARIMemoryDeleter del,
                             const void *user, ANARIDataType t, uint64_t n1) {
  return reinterpret_cast<ANARIArray1D>(new_array(d, app, del, user, t, n1, 1, 1, ANARI_ARRAY1D));
}
ANARIArray2D anariNewArray2D(ANARIDevice d, const void *app, ANARIMemoryDeleter del,
                             const void *user, ANARIDataType t, uint64_t n1, uint64_t n2) {
  return reinterpret_cast<ANARIArray2D>(new_array(d, app, del, user, t, n1, n2, 1, ANARI_ARRAY2D));
}
ANARIArray3D anariNewArray3D(ANARIDevice d, const void *app, ANARIMemoryDeleter del,
                             const void *user, ANARIDataType t, uint64_t n1, uint64_t n2,
                             uint64_t n3) {
  return reinterpret_cast<ANARIArray3D>(
      new_array(d, app, del, user, t, n1, n2, n3, ANARI_ARRAY3D));
}

void *anariMapArray(ANARIDevice, ANARIArray ah) {
  Array *a = reinterpret_cast<Array *>(ah);
  if (!a) return nullptr;
  a->mapped = true;
  return const_cast<uint8_t *>(a->data());
}
void anariUnmapArray(ANARIDevice dh, ANARIArray ah) {
  Array *a = reinterpret_cast<Array *>(ah);
  if (!a || !a->mapped) return;
  a->mapped = false;
  a->hold_elements();
  a->version = ++as_device(dh)->serial;  // its users see new contents
}

#define VRT_NEW(fn, T, kind, sub)                                   \
  {                                                                 \
    Device *d = as_device(dh);                                      \
    if (!d) return nullptr;                                         \
    return reinterpret_cast<ANARIObject>(make<T>(d, kind, sub));    \
  }

ANARIGeometry anariNewGeometry(ANARIDevice dh, const char *t) VRT_NEW(g, Geometry, ANARI_GEOMETRY, t)
ANARIMaterial anariNewMaterial(ANARIDevice dh, const char *t) VRT_NEW(m, Object, ANARI_MATERIAL, t)
ANARISurface anariNewSurface(ANARIDevice dh) VRT_NEW(s, Object, ANARI_SURFACE, "")
ANARISpatialField anariNewSpatialField(ANARIDevice dh, const char *t)
    VRT_NEW(f, Field, ANARI_SPATIAL_FIELD, t)
ANARIVolume anariNewVolume(ANARIDevice dh, const char *t) VRT_NEW(v, Volume, ANARI_VOLUME, t)
ANARILight anariNewLight(ANARIDevice dh, const char *t) VRT_NEW(l, Light, ANARI_LIGHT, t)
ANARIGroup anariNewGroup(ANARIDevice dh) VRT_NEW(g, Object, ANARI_GROUP, "")
ANARIInstance anariNewInstance(ANARIDevice dh, const char *t) VRT_NEW(i, Object, ANARI_INSTANCE, t)
ANARIWorld anariNewWorld(ANARIDevice dh) VRT_NEW(w, Object, ANARI_WORLD, "")
ANARICamera anariNewCamera(ANARIDevice dh, const char *t) VRT_NEW(c, Object, ANARI_CAMERA, t)
ANARIRenderer anariNewRenderer(ANARIDevice dh, const char *t) VRT_NEW(r, Object, ANARI_RENDERER, t)
ANARIFrame anariNewFrame(ANARIDevice dh) VRT_NEW(f, Frame, ANARI_FRAME, "")

#undef VRT_NEW

void anariSetParameter(ANARIDevice dh, ANARIObject oh, const char *name, ANARIDataType t,
                       const void *mem) {
  Device *d = as_device(dh);
  Object *o = as_obj(oh);
  if (!d || !o || !name || !mem) return;
  Param p;
  p.type = t;
  if (t == ANARI_STRING) {
    p.str = static_cast<const char *>(mem);
  } else if (is_object_type(t)) {
    p.obj = *static_cast<Object *const *>(mem);
    retain_internal(p.obj);
  } else {
    const size_t n = type_size(t);
    if (n == 0) {
      d->report(o, ANARI_SEVERITY_WARNING, ANARI_STATUS_INVALID_ARGUMENT,
                "parameter '%s' of unsupported type %d ignored", name, t);
      return;
    }
    p.bytes.assign(static_cast<const uint8_t *>(mem), static_cast<const uint8_t *>(mem) + n);
  }
  auto it = o->staged.find(name);
  if (it != o->staged.end() && it->second.obj) release_internal(it->second.obj);
  o->staged[name] = p;
}

void anariUnsetParameter(ANARIDevice, ANARIObject oh, const char *name) {
  Object *o = as_obj(oh);
  if (!o || !name) return;
  auto it = o->staged.find(name);
  if (it == o->staged.end()) return;
  if (it->second.obj) release_internal(it->second.obj);
  o->staged.erase(it);
}

void anariUnsetAllParameters(ANARIDevice, ANARIObject oh) {
  Object *o = as_obj(oh);
  if (!o) return;
  for (auto &kv : o->staged)
    if (kv.second.obj) release_internal(kv.second.obj);
  o->staged.clear();
}

void anariCommitParameters(ANARIDevice dh, ANARIObject oh) {
  Device *d = as_device(dh);
  Object *o = as_obj(oh);
  if (!d || !o) return;
  if (o == d) {
    d->flush();  // a device commit flushes the commit buffer
    return;
  }
  if (o->pending) return;
  o->pending = true;
  retain_internal(o);  // the buffer keeps the object alive until the flush
  d->commit_buffer.push_back(o);
}

void anariRelease(ANARIDevice, ANARIObject oh) {
  Object *o = as_obj(oh);
  if (!o) return;
  if (o->refs.fetch_sub(1) == 1) destroy(o);
}

void anariRetain(ANARIDevice, ANARIObject oh) {
  if (Object *o = as_obj(oh)) o->refs.fetch_add(1);
}

int anariGetProperty(ANARIDevice dh, ANARIObject oh, const char *name, ANARIDataType t,
                     void *mem, uint64_t size, ANARIWaitMask mask) {
  Device *d = as_device(dh);
  Object *o = as_obj(oh);
  if (!d || !o || !name || !mem) return 0;
  if (o->kind == ANARI_GEOMETRY) {
    // extension: the triangle BVH of the last build ("bvhBuildTime" in
    // seconds, "bvhNodes"); the build itself happens when a frame uses the
    // geometry after its commit
    d->flush();
    const vrt_bvh b = static_cast<Geometry *>(o)->bvh;
    vrt_bvh_info info{};
    if (!b || vrt_bvh_get_info(b, &info) != VRT_OK) return 0;
    if (!std::strcmp(name, "bvhBuildTime") && t == ANARI_FLOAT32 && size >= 4) {
      const float sec = float(info.build_ms * 1e-3);
      std::memcpy(mem, &sec, 4);
      return 1;
    }
    if (!std::strcmp(name, "bvhNodes") && (t == ANARI_INT32 || t == ANARI_UINT32) && size >= 4) {
      const uint32_t n = info.num_nodes;
      std::memcpy(mem, &n, 4);
      return 1;
    }
    return 0;
  }
  if (o->kind != ANARI_FRAME) return 0;
  Frame *fr = static_cast<Frame *>(o);
  if (mask == ANARI_WAIT) (void)hipStreamSynchronize(d->stream);
  if (!std::strcmp(name, "duration") && t == ANARI_FLOAT32 && size >= 4) {
    if (!fr->timed || hipEventQuery(fr->e1) != hipSuccess) return 0;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, fr->e0, fr->e1) != hipSuccess) return 0;
    const float s = ms * 1e-3f;
    std::memcpy(mem, &s, 4);
    return 1;
  }
  if (!std::strcmp(name, "numSamples") && (t == ANARI_INT32 || t == ANARI_UINT32) && size >= 4) {
    const int32_t n = fr->accum_id;
    std::memcpy(mem, &n, 4);
    return 1;
  }
  return 0;
}

void anariRenderFrame(ANARIDevice dh, ANARIFrame fh) {
  Device *d = as_device(dh);
  Object *o = as_obj(fh);
  if (!d || !o || o->kind != ANARI_FRAME) return;
  render(d, static_cast<Frame *>(o));
}

int anariFrameReady(ANARIDevice dh, ANARIFrame, ANARIWaitMask mask) {
  Device *d = as_device(dh);
  if (!d) return 0;
  d->flush();
  if (mask == ANARI_WAIT) return hipStreamSynchronize(d->stream) == hipSuccess ? 1 : 0;
  return hipStreamQuery(d->stream) == hipSuccess ? 1 : 0;
}

void anariDiscardFrame(ANARIDevice, ANARIFrame) {}

const void *anariMapFrame(ANARIDevice dh, ANARIFrame fh, const char *channel, uint32_t *w,
                          uint32_t *h, ANARIDataType *type) {
  Device *d = as_device(dh);
  Object *o = as_obj(fh);
  if (w) *w = 0;
  if (h) *h = 0;
  if (type) *type = ANARI_UNKNOWN;
  if (!d || !o || o->kind != ANARI_FRAME || !channel) return nullptr;
  Frame *fr = static_cast<Frame *>(o);
  if (!fr->color.p) return nullptr;
  (void)hipStreamSynchronize(d->stream);
  const size_t np = size_t(fr->w) * fr->h;
  if (!std::strcmp(channel, "channel.color")) {
    const ANARIDataType ct = fr->color_type == ANARI_UNKNOWN ? ANARI_FLOAT32_VEC4 : fr->color_type;
    fr->host_color.resize(np * (ct == ANARI_FLOAT32_VEC4 ? 16 : 4));
    if (vrt_download(d->ctx, fr->host_color.data(), fr->color.p, fr->host_color.size()) != VRT_OK) {
      d->vrt_error(fr, "map channel.color");
      return nullptr;
    }
    if (w) *w = fr->w;
    if (h) *h = fr->h;
    if (type) *type = ct;
    return fr->host_color.data();
  }
  if (!std::strcmp(channel, "channel.depth") && fr->depth.p) {
    fr->host_depth.resize(np);
    if (vrt_download(d->ctx, fr->host_depth.data(), fr->depth.p, np * 4) != VRT_OK) {
      d->vrt_error(fr, "map channel.depth");
      return nullptr;
    }
    if (w) *w = fr->w;
    if (h) *h = fr->h;
    if (type) *type = ANARI_FLOAT32;
    return fr->host_depth.data();
  }
  for (int k = 0; k < 5; ++k) {
    if (std::strcmp(channel, kGbuf[k].name) || !fr->gbuf[k].p) continue;
    fr->host_gbuf[k].resize(np * kGbuf[k].bytes);
    if (vrt_download(d->ctx, fr->host_gbuf[k].data(), fr->gbuf[k].p, fr->host_gbuf[k].size()) !=
        VRT_OK) {
      d->vrt_error(fr, "map g-buffer channel");
      return nullptr;
    }
    if (w) *w = fr->w;
    if (h) *h = fr->h;
    if (type) *type = kGbuf[k].type;
    return fr->host_gbuf[k].data();
  }
  d->report(fr, ANARI_SEVERITY_WARNING, ANARI_STATUS_INVALID_ARGUMENT,
            "frame has no channel '%s'", channel);
  return nullptr;
}

void anariUnmapFrame(ANARIDevice, ANARIFrame, const char *) {}

}  // extern "C"
