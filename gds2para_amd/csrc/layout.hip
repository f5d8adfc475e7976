// layout.hip -- gp_layout lifetime: validation, device upload of the mesh
// contract, and the PEC edge map (setMapEdge, src/fdtd.hpp:1508-1529) built on
// the device with a scan over the removed-edge flags.
#include <hipcub/hipcub.hpp>

#include <cstring>

#include "gp_dev.hpp"

namespace gp {

void HostMesh::copy_from(const gp_mesh &m) {
  Grid g = Grid::make(m.nx, m.ny, m.nz);
  xn.assign(m.xn, m.xn + m.nx);
  yn.assign(m.yn, m.yn + m.ny);
  zn.assign(m.zn, m.zn + m.nz);
  epsn.assign(m.stack_epsn, m.stack_epsn + m.nz);
  sign.assign(m.stack_sign, m.stack_sign + m.nz);
  mark_edge.assign(m.mark_edge, m.mark_edge + g.Ne);
  mark_node.assign(m.mark_node, m.mark_node + g.Nn);
  cdt_mark_port.assign(m.cdt_mark_port, m.cdt_mark_port + m.num_cdt);
  bd_edges.assign(m.bd_edges, m.bd_edges + m.num_bd_edges);
  bd_nodes.assign(m.bd_nodes, m.bd_nodes + m.num_bd_nodes);
  int32_t ns = m.num_ports > 0 ? m.port_side_ptr[m.num_ports] : 0;
  if (m.num_ports > 0) {
    port_side_ptr.assign(m.port_side_ptr, m.port_side_ptr + m.num_ports + 1);
    side_edge_ptr.assign(m.side_edge_ptr, m.side_edge_ptr + ns + 1);
    side_edges.assign(m.side_edges, m.side_edges + m.side_edge_ptr[ns]);
    side_area.assign(m.side_area, m.side_area + ns);
    side_dir.assign(m.side_dir, m.side_dir + ns);
    side_cnd.assign(m.side_cnd, m.side_cnd + ns);
  } else {
    port_side_ptr = {0};
    side_edge_ptr = {0};
  }
  view = m;
  view.xn = xn.data(); view.yn = yn.data(); view.zn = zn.data();
  view.stack_epsn = epsn.data(); view.stack_sign = sign.data();
  view.mark_edge = mark_edge.data(); view.mark_node = mark_node.data();
  view.cdt_mark_port = cdt_mark_port.data();
  view.bd_edges = bd_edges.data(); view.bd_nodes = bd_nodes.data();
  view.port_side_ptr = port_side_ptr.data(); view.side_edge_ptr = side_edge_ptr.data();
  view.side_edges = side_edges.data(); view.side_area = side_area.data();
  view.side_dir = side_dir.data(); view.side_cnd = side_cnd.data();
}

gp_status ensure_device() {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n < 1) {
    (void)hipGetLastError();
    return fail(GP_ERR_NODEVICE, "no HIP device visible");
  }
  return GP_OK;
}

gp_status exclusive_scan_i64(const int64_t *cnt, int64_t *off, int64_t n, hipStream_t s, Arena &tmp) {
  // off[0..n] = exclusive prefix of cnt; off[n] = total (cnt must hold n entries)
  GP_HIP(hipMemsetAsync(off, 0, sizeof(int64_t), s));
  if (n == 0) return GP_OK;
  size_t bytes = 0;
  GP_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, bytes, cnt, off + 1, (int)0, s));
  bytes = 0;
  // hipcub takes an int item count; split very long scans is unnecessary here (n < 2^31)
  if (n > 2147483647LL) return fail(GP_ERR_ARG, "scan longer than 2^31 items");
  GP_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, bytes, cnt, off + 1, (int)n, s));
  uint8_t *buf = nullptr;
  GP_TRY(tmp.alloc(&buf, (int64_t)bytes + 1));
  GP_HIP(hipcub::DeviceScan::InclusiveSum(buf, bytes, cnt, off + 1, (int)n, s));
  return GP_OK;
}

}  // namespace gp

using namespace gp;

namespace {

__global__ void k_flags(const int64_t *__restrict__ idx, int64_t n, uint8_t *__restrict__ flag) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flag[idx[i]] = 1;
}

__global__ void k_removed_count(const uint8_t *__restrict__ rm, int64_t ne, int64_t *__restrict__ c) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < ne) c[e] = rm[e];
}

// mapEdge[e] = e - (#removed before e) or -1; mapEdgeR is its inverse.
__global__ void k_map_edge(const uint8_t *__restrict__ rm, const int64_t *__restrict__ before, int64_t ne,
                           int64_t *__restrict__ map, int64_t *__restrict__ mapr) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  if (rm[e]) { map[e] = -1; return; }
  int64_t m = e - before[e];
  map[e] = m;
  mapr[m] = e;
}

}  // namespace

extern "C" gp_status gp_device_query(gp_device_info *out) {
  if (!out) return fail(GP_ERR_ARG, "null argument");
  std::memset(out, 0, sizeof(*out));
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) { (void)hipGetLastError(); n = 0; }
  out->device_count = n;
  if (n < 1) return fail(GP_ERR_NODEVICE, "no HIP device visible");
  hipDeviceProp_t p;
  GP_HIP(hipGetDeviceProperties(&p, 0));
  std::strncpy(out->arch, p.gcnArchName, sizeof(out->arch) - 1);
  out->total_mem = (int64_t)p.totalGlobalMem;
  out->cu_count = p.multiProcessorCount;
  return GP_OK;
}

extern "C" gp_status gp_layout_create(const gp_mesh *mesh, gp_layout **out) {
  if (!out) return fail(GP_ERR_ARG, "null argument");
  *out = nullptr;
  GP_TRY(gp_mesh_validate(mesh));
  GP_TRY(ensure_device());
  auto lay = std::make_unique<gp_layout>();
  lay->hm.copy_from(*mesh);
  const gp_mesh &m = lay->hm.view;
  lay->g = Grid::make(m.nx, m.ny, m.nz);
  const Grid &g = lay->g;
  lay->n = g.Ne - m.num_bd_edges;
  GP_HIP(hipGetDevice(&lay->device));
  GP_HIP(hipStreamCreateWithFlags(&lay->stream, hipStreamNonBlocking));
  hipStream_t s = lay->stream;
  Arena &A = lay->arena;
  GP_TRY(A.alloc(&lay->xn, m.nx)); GP_TRY(A.alloc(&lay->yn, m.ny)); GP_TRY(A.alloc(&lay->zn, m.nz));
  GP_TRY(A.alloc(&lay->epsn, m.nz)); GP_TRY(A.alloc(&lay->sign, m.nz));
  GP_TRY(A.alloc(&lay->mark_edge, g.Ne)); GP_TRY(A.alloc(&lay->mark_node, g.Nn));
  GP_TRY(A.alloc(&lay->rm_edge, g.Ne)); GP_TRY(A.alloc(&lay->bd_node, g.Nn));
  GP_TRY(A.alloc(&lay->map_edge, g.Ne)); GP_TRY(A.alloc(&lay->map_edge_r, lay->n));
  GP_HIP(hipMemcpyAsync(lay->xn, m.xn, sizeof(double) * m.nx, hipMemcpyHostToDevice, s));
  GP_HIP(hipMemcpyAsync(lay->yn, m.yn, sizeof(double) * m.ny, hipMemcpyHostToDevice, s));
  GP_HIP(hipMemcpyAsync(lay->zn, m.zn, sizeof(double) * m.nz, hipMemcpyHostToDevice, s));
  GP_HIP(hipMemcpyAsync(lay->epsn, m.stack_epsn, sizeof(double) * m.nz, hipMemcpyHostToDevice, s));
  GP_HIP(hipMemcpyAsync(lay->sign, m.stack_sign, sizeof(double) * m.nz, hipMemcpyHostToDevice, s));
  GP_HIP(hipMemcpyAsync(lay->mark_edge, m.mark_edge, sizeof(int32_t) * g.Ne, hipMemcpyHostToDevice, s));
  GP_HIP(hipMemcpyAsync(lay->mark_node, m.mark_node, sizeof(int32_t) * g.Nn, hipMemcpyHostToDevice, s));
  GP_HIP(hipMemsetAsync(lay->rm_edge, 0, g.Ne, s));
  GP_HIP(hipMemsetAsync(lay->bd_node, 0, g.Nn, s));
  Arena tmp;
  int64_t *d_idx = nullptr, *cnt = nullptr, *before = nullptr;
  int64_t nb = std::max<int64_t>(m.num_bd_edges, m.num_bd_nodes);
  gp_status st = GP_OK;
  do {
    if ((st = tmp.alloc(&d_idx, std::max<int64_t>(nb, 1))) != GP_OK) break;
    if ((st = tmp.alloc(&cnt, g.Ne)) != GP_OK) break;
    if ((st = tmp.alloc(&before, g.Ne + 1)) != GP_OK) break;
    if (m.num_bd_edges > 0) {
      if (hipMemcpyAsync(d_idx, m.bd_edges, sizeof(int64_t) * m.num_bd_edges, hipMemcpyHostToDevice, s) != hipSuccess) { st = fail(GP_ERR_HIP, "copy bd_edges"); break; }
      k_flags<<<blocks_for(m.num_bd_edges), 256, 0, s>>>(d_idx, m.num_bd_edges, lay->rm_edge);
    }
    k_removed_count<<<blocks_for(g.Ne), 256, 0, s>>>(lay->rm_edge, g.Ne, cnt);
    if ((st = exclusive_scan_i64(cnt, before, g.Ne, s, tmp)) != GP_OK) break;
    k_map_edge<<<blocks_for(g.Ne), 256, 0, s>>>(lay->rm_edge, before, g.Ne, lay->map_edge, lay->map_edge_r);
    if (m.num_bd_nodes > 0) {
      if (hipStreamSynchronize(s) != hipSuccess) { st = fail(GP_ERR_HIP, "sync"); break; }
      if (hipMemcpyAsync(d_idx, m.bd_nodes, sizeof(int64_t) * m.num_bd_nodes, hipMemcpyHostToDevice, s) != hipSuccess) { st = fail(GP_ERR_HIP, "copy bd_nodes"); break; }
      k_flags<<<blocks_for(m.num_bd_nodes), 256, 0, s>>>(d_idx, m.num_bd_nodes, lay->bd_node);
    }
    if (hipStreamSynchronize(s) != hipSuccess || hipGetLastError() != hipSuccess) { st = fail(GP_ERR_HIP, "layout upload kernels"); break; }
  } while (0);
  (void)hipStreamSynchronize(s);
  tmp.free_all();
  if (st != GP_OK) {
    lay->arena.free_all();
    (void)hipStreamDestroy(lay->stream);
    return st;
  }
  *out = lay.release();
  return GP_OK;
}

extern "C" void gp_layout_destroy(gp_layout *lay) {
  if (!lay) return;
  if (lay->stream) (void)hipStreamSynchronize(lay->stream);
  destroy_solver(lay->solver);
  lay->solver = nullptr;
  destroy_v0path(lay->v0path);
  lay->v0path = nullptr;
  destroy_v0(lay->v0);
  lay->v0 = nullptr;
  lay->arena.free_all();
  if (lay->stream) (void)hipStreamDestroy(lay->stream);
  delete lay;
}

extern "C" int64_t gp_layout_device_bytes(const gp_layout *lay) {
  return lay ? lay->arena.bytes + v0path_bytes(lay->v0path) + solver_bytes(lay->solver) : 0;
}
