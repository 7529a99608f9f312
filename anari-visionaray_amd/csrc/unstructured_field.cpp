// vrt_unstructured: a tetrahedral field (UnstructuredField.cpp:17-198)
// converted once into the GPU point-location layout of tet_grid.h.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

#include "tet_grid.h"
#include "unstructured_internal.h"
#include "vrt.h"

namespace {

using namespace vrt;

// cell.type codes (UnstructuredField.cpp:82-104): barney and VTK numbering
enum class Elem { Tet, Hex, Wedge, Pyr, BezierHex, Unknown };

Elem map_type(uint8_t t) {
  if (t == 0 || t == 10) return Elem::Tet;
  if (t == 1 || t == 12) return Elem::Hex;
  if (t == 2 || t == 13) return Elem::Wedge;
  if (t == 3 || t == 14) return Elem::Pyr;
  if (t == 79) return Elem::BezierHex;
  return Elem::Unknown;
}

}  // namespace

extern "C" {

int vrt_unstructured_create(vrt_context ctx, const float *pos, const float *vdata, int64_t nv,
                            const uint32_t *index, int64_t ni, const uint32_t *cell_index,
                            const uint8_t *cell_type, const float *cdata, int64_t nc,
                            vrt_unstructured *out) {
  if (int r = check_ctx(ctx)) return r;
  if (!out) return fail(VRT_ERROR_INVALID_ARGUMENT, "null output handle");
  *out = nullptr;
  if (!pos || nv < 1) return fail(VRT_ERROR_INVALID_ARGUMENT, "missing vertex.position");
  if (!vdata && !cdata) return fail(VRT_ERROR_INVALID_ARGUMENT, "missing vertex.data and cell.data");
  if (!index || ni < 1) return fail(VRT_ERROR_INVALID_ARGUMENT, "missing index");
  if (!cell_index || !cell_type || nc < 1)
    return fail(VRT_ERROR_INVALID_ARGUMENT, "missing cell.index or cell.type");
  if (nv > 0xFFFFFFFFll || ni > 0xFFFFFFFFll || nc > 0x7FFFFFFFll)
    return fail(VRT_ERROR_UNSUPPORTED, "field exceeds 32-bit indexing");
  const auto t0 = std::chrono::steady_clock::now();

  std::vector<tetgrid::Rec> recs;
  std::vector<float4> tets;  // {bbox min, value min}, {bbox max, value max}
  std::vector<uint32_t> cell_of;  // kept tetrahedron -> ANARI cell index
  recs.reserve(size_t(nc));
  cell_of.reserve(size_t(nc));
  tets.reserve(2 * size_t(nc));
  uint64_t skipped = 0;
  float blo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, bhi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
  float avg_diag = 0.f;
  const float nan = std::numeric_limits<float>::quiet_NaN();
  for (int64_t c = 0; c < nc; ++c) {
    const uint64_t first = cell_index[c];
    const uint64_t last = c < nc - 1 ? uint64_t(cell_index[c + 1]) : uint64_t(ni);
    if (last < first || last > uint64_t(ni))
      return fail(VRT_ERROR_INVALID_ARGUMENT, "cell.index of cell %lld out of order or range",
                  (long long)c);
    const uint64_t ic = last - first;
    const Elem type = map_type(cell_type[c]);
    // skipped with a warning by the reference (UnstructuredField.cpp:135-147)
    if (ic < 4 || (type == Elem::Tet && ic != 4) || (type == Elem::Pyr && ic != 5) ||
        (type == Elem::Wedge && ic != 6) || (type == Elem::Hex && ic != 8)) {
      ++skipped;
      continue;
    }
    // the reference intersects by vertex count (DeviceCopyableObjects.h:145-148)
    if (ic != 4)
      return fail(VRT_ERROR_UNSUPPORTED,
                  "cell %lld has %llu vertices: only tetrahedral cells are supported",
                  (long long)c, (unsigned long long)ic);
    tetgrid::Rec r;
    const float cell_value = cdata ? cdata[c] : 0.f;
    float bmn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, bmx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    float vmn = FLT_MAX, vmx = -FLT_MAX;
    for (int k = 0; k < 4; ++k) {
      const uint32_t i = index[first + k];
      if (int64_t(i) >= nv)
        return fail(VRT_ERROR_INVALID_ARGUMENT, "index %u of cell %lld outside the vertices", i,
                    (long long)c);
      for (int a = 0; a < 3; ++a) {
        r.v[k][a] = pos[3 * size_t(i) + a];
        bmn[a] = std::fmin(bmn[a], r.v[k][a]);
        bmx[a] = std::fmax(bmx[a], r.v[k][a]);
      }
      r.v[k][3] = vdata ? vdata[i] : nan;
      // buildGrid's value range (UnstructuredField.cpp:387-399)
      const float w = std::isnan(r.v[k][3]) ? cell_value : r.v[k][3];
      vmn = std::fmin(vmn, w);
      vmx = std::fmax(vmx, w);
    }
    // no data at the first vertex: the cell value (DeviceCopyableObjects.h:150-153)
    if (std::isnan(r.v[0][3])) r.v[1][3] = cell_value;
    recs.push_back(r);
    cell_of.push_back(uint32_t(c));
    tets.push_back(make_float4(bmn[0], bmn[1], bmn[2], vmn));
    tets.push_back(make_float4(bmx[0], bmx[1], bmx[2], vmx));
    for (int a = 0; a < 3; ++a) {
      blo[a] = std::fmin(blo[a], bmn[a]);
      bhi[a] = std::fmax(bhi[a], bmx[a]);
    }
    // avgCellDiagonal as the reference accumulates it: each diagonal divided
    // by the number of cells kept so far (UnstructuredField.cpp:170)
    const float dx = bmx[0] - bmn[0], dy = bmx[1] - bmn[1], dz = bmx[2] - bmn[2];
    avg_diag += std::sqrt(dx * dx + dy * dy + dz * dz) / float(recs.size());
  }
  if (recs.empty()) return fail(VRT_ERROR_INVALID_ARGUMENT, "no valid cells");

  tetgrid::Grid g;
  if (!tetgrid::build(recs, blo, bhi, 1.0, 0, g))
    return fail(VRT_ERROR_UNSUPPORTED, "point-location lists exceed 2^32 entries");
  const auto t1 = std::chrono::steady_clock::now();

  auto *h = new vrt_unstructured_s;
  h->device = ctx->device;
  for (int a = 0; a < 3; ++a) {
    h->lo[a] = g.lo[a];
    h->inv_cell[a] = g.inv_cell[a];
    h->dims[a] = g.dims[a];
    h->blo[a] = blo[a];
    h->bhi[a] = bhi[a];
  }
  h->cell_size = avg_diag;
  h->num_tets = recs.size();
  h->skipped = skipped;
  h->entries = g.recs.size();
  h->max_list = g.max_list;
  h->mean_list = g.mean_list;
  h->expected_tests = g.expected_tests;
  h->build_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  auto cleanup = [&](hipError_t e, const char *what) {
    (void)hipFree(h->start);
    (void)hipFree(h->recs);
    (void)hipFree(h->rec_cell);
    (void)hipFree(h->tets);
    delete h;
    return hip_fail(e, what);
  };
  // the primitive id of a list entry (UnstructuredField.cpp:146-151: elemID =
  // cell index), read only for the g-buffer
  std::vector<uint32_t> rec_cell(g.rec_tet.size());
  for (size_t i = 0; i < rec_cell.size(); ++i) rec_cell[i] = cell_of[g.rec_tet[i]];
  std::vector<uint32_t>().swap(g.rec_tet);
  const size_t start_bytes = g.start.size() * 4, rec_bytes = g.recs.size() * sizeof(tetgrid::Rec),
               cell_bytes = rec_cell.size() * 4, tet_bytes = tets.size() * sizeof(float4);
  hipError_t e = hipMalloc(&h->start, start_bytes);
  if (e == hipSuccess) e = hipMalloc(&h->recs, std::max<size_t>(rec_bytes, 64));
  if (e == hipSuccess) e = hipMalloc(&h->rec_cell, std::max<size_t>(cell_bytes, 4));
  if (e == hipSuccess) e = hipMalloc(&h->tets, tet_bytes);
  if (e != hipSuccess) return cleanup(e, "unstructured field allocation");
  e = hipMemcpyAsync(h->start, g.start.data(), start_bytes, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess && rec_bytes)
    e = hipMemcpyAsync(h->recs, g.recs.data(), rec_bytes, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess && cell_bytes)
    e = hipMemcpyAsync(h->rec_cell, rec_cell.data(), cell_bytes, hipMemcpyHostToDevice,
                       ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(h->tets, tets.data(), tet_bytes, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);  // host copies go out of scope
  if (e != hipSuccess) return cleanup(e, "unstructured field upload");
  h->device_bytes = start_bytes + rec_bytes + cell_bytes + tet_bytes;
  *out = h;
  return VRT_OK;
}

int vrt_unstructured_get_info(vrt_unstructured f, vrt_unstructured_info *info) {
  if (!f || !info) return fail(VRT_ERROR_INVALID_ARGUMENT, "null argument");
  for (int a = 0; a < 3; ++a) {
    info->bounds_lo[a] = f->blo[a];
    info->bounds_hi[a] = f->bhi[a];
    info->grid_dims[a] = f->dims[a];
  }
  info->cell_size = f->cell_size;
  info->num_cells = f->num_tets;
  info->skipped_cells = f->skipped;
  info->list_entries = f->entries;
  info->max_list = f->max_list;
  info->mean_list = f->mean_list;
  info->expected_tests = f->expected_tests;
  info->build_ms = f->build_ms;
  info->device_bytes = f->device_bytes;
  return VRT_OK;
}

// Frees the device copies; the creating context may already be gone, so
// only the device id is used (hipFree waits for work that reads them).
int vrt_unstructured_destroy(vrt_unstructured f) {
  if (!f) return VRT_OK;
  (void)hipSetDevice(f->device);
  (void)hipFree(f->start);
  (void)hipFree(f->recs);
  (void)hipFree(f->rec_cell);
  (void)hipFree(f->tets);
  delete f;
  return VRT_OK;
}

}  // extern "C"
