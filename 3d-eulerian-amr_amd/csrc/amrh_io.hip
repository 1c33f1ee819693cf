// Checkpoints in the reference's on-disk layout (GridNode::write / read, grid/grid_node.cpp:29-57,
// over Grid::write / read, grid/grid.cpp:104-146), so a run can move between this library and the
// reference.  Depth first from the root, children in index order; per node:
//   double h, double h0, double time, 3 x (int offset, double origin), double FO[nf],
//   the padded GNX^3 state with x slowest and the field fastest ((*this)(i, j, k)[l] for i, j, k, l),
//   8 x bool child-present, double refine_alpha, int age.
// Edge and corner ghost cells are not held by the device; they are written as zeros and ignored
// on reading (the reference never reads them, grid/grid_node.cpp:59-112).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>

#include "amrh_ctx.h"
#include "amrh_internal.h"
#include "amrhydro.h"

using namespace amrh;

namespace {

struct Writer {
  FILE* fp;
  bool ok = true;
  template <typename T>
  void put(const T& v) {
    ok = ok && fwrite(&v, sizeof(T), 1, fp) == 1;
  }
};

struct Reader {
  FILE* fp;
  bool ok = true;
  template <typename T>
  T get() {
    T v{};
    ok = ok && fread(&v, sizeof(T), 1, fp) == 1;
    return v;
  }
};

// GridNode::output's zones of one node (grid/grid_node.cpp:1326-1346): interior cells in k, j, i
// order, skipping those covered by a child; a zone's rank among its node's zones in closed form.
template <int N, int G>
__global__ __launch_bounds__(256) void zone_gather_kernel(const double* U, int nf, const int* nodes,
                                                          const long* zone0, const unsigned char* covered,
                                                          long nzones, double* out) {
  constexpr long R = (long)N * N * N + 6L * G * N * N;
  constexpr int H = N / 2;
  const int b = nodes[blockIdx.x];
  const unsigned cov = covered[b];
  auto open = [&](int cx, int cy, int cz) { return !((cov >> (cx | (cy << 1) | (cz << 2))) & 1u); };
  const double* Ub = U + (long)b * nf * R;
  for (int e = threadIdx.x; e < N * N * N; e += blockDim.x) {
    const int i = e % N, j = (e / N) % N, k = e / (N * N);
    const int cx = i / H, cy = j / H, cz = k / H;
    if (!open(cx, cy, cz)) continue;
    long r = 0;
    for (int z = 0; z < cz; z++)  // whole planes below
      r += (long)H * H * H * ((int)open(0, 0, z) + open(1, 0, z) + open(0, 1, z) + open(1, 1, z));
    r += (long)(k - cz * H) * H * H * ((int)open(0, 0, cz) + open(1, 0, cz) + open(0, 1, cz) + open(1, 1, cz));
    for (int y = 0; y < cy; y++) r += (long)H * H * ((int)open(0, y, cz) + open(1, y, cz));
    r += (long)(j - cy * H) * H * ((int)open(0, cy, cz) + open(1, cy, cz));
    if (cx == 1 && open(0, cy, cz)) r += H;
    r += i - cx * H;
    const long z = zone0[blockIdx.x] + r;
    for (int l = 0; l < nf; l++) out[(long)l * nzones + z] = Ub[l * R + e];
  }
}

// the nodes in depth-first pre-order (GridNode::output's recursion, grid/grid_node.cpp:1348-1352)
std::vector<int> preorder(const TreeState* t) {
  std::vector<int> out, stack = {0};
  while (!stack.empty()) {
    const int b = stack.back();
    stack.pop_back();
    out.push_back(b);
    for (int c = 7; c >= 0; c--)
      if (t->child[b][c] >= 0) stack.push_back(t->child[b][c]);
  }
  return out;
}

int open_octants(const TreeState* t, int b) {
  int o = 0;
  for (int c = 0; c < 8; c++) o += t->child[b][c] < 0;
  return o;
}

void put_be(FILE* fp, const void* v, int bytes, bool* ok) {  // legacy VTK binary is big-endian
  unsigned char buf[8];
  const unsigned char* p = (const unsigned char*)v;
  for (int q = 0; q < bytes; q++) buf[q] = p[bytes - 1 - q];
  *ok = *ok && fwrite(buf, 1, bytes, fp) == (size_t)bytes;
}

}  // namespace

extern "C" {

int amrh_tree_output_counts(amrh_ctx* c, long* npoints, long* nzones) {
  if (!c || !npoints || !nzones) return fail(AMRH_EINVAL, "null argument");
  if (!c->tree) return fail(AMRH_ESTATE, "output needs an octree: call amrh_tree_set");
  const long G1 = c->cfg.n + 2 * c->cfg.g + 1, H = c->cfg.n / 2;
  long z = 0;
  for (int b = 0; b < c->nb; b++) z += open_octants(c->tree, b) * H * H * H;
  *npoints = (long)c->nb * G1 * G1 * G1;  // GRID_NNODES per node (grid/grid.h:11)
  *nzones = z;
  return AMRH_OK;
}

int amrh_tree_output_mesh(amrh_ctx* c, double* x, double* y, double* z, int* nodelist, double* values) {
  if (!c) return fail(AMRH_EINVAL, "null argument");
  long npts, nz;
  if (int rc = amrh_tree_output_counts(c, &npts, &nz)) return rc;
  if (!c->have_u) return fail(AMRH_ESTATE, "no state uploaded");
  if (!((c->cfg.g == 2 && (c->cfg.n == 8 || c->cfg.n == 16)) || (c->cfg.g == 3 && c->cfg.n == 8)))
    return fail(AMRH_EINVAL, "n = 8 or 16 with g = 2, or n = 8 with g = 3");
  const TreeState* t = c->tree;
  const int n = c->cfg.n, g = c->cfg.g, GNX = n + 2 * g, G1 = GNX + 1, H = n / 2;
  const std::vector<int> order = preorder(t);
  const double h0 = t->h[0], gd = c->cfg.grid_dim;
  long pi = 0, ni = 0;
  std::vector<long> zone0(order.size());
  long zacc = 0;
  for (size_t q = 0; q < order.size(); q++) {
    const int b = order[q];
    const double h = t->h[b];
    zone0[q] = zacc;
    zacc += open_octants(t, b) * (long)H * H * H;
    // Grid::xf / yf / zf (grid/grid.cpp:510-526); z subtracts the origin only under Z_REFLECT
    if (x && y && z) {
      const double* o = c->origin;
      for (int k = 0; k < G1; k++)
        for (int j = 0; j < G1; j++)
          for (int i = 0; i < G1; i++) {
            const long p = pi + ((long)k * G1 + j) * G1 + i;
            x[p] = double(t->off[b][0] + i) * h - gd - double(g) * h0 - o[0];
            y[p] = double(t->off[b][1] + j) * h - gd - double(g) * h0 - o[1];
            z[p] = c->cfg.z_reflect ? double(t->off[b][2] + k) * h - double(g) * h0 - o[2]
                                    : double(t->off[b][2] + k) * h - gd - double(g) * h0;
          }
    }
    if (nodelist) {
      for (int k = g; k < GNX - g; k++)
        for (int j = g; j < GNX - g; j++)
          for (int i = g; i < GNX - g; i++) {
            const int cidx = ((2 * i) / GNX) | (((2 * j) / GNX) << 1) | (((2 * k) / GNX) << 2);
            if (t->child[b][cidx] >= 0) continue;  // GridNode::zone_is_refined
            auto at = [&](int ii, int jj, int kk) { return (int)(ii + G1 * (jj + G1 * kk) + pi); };
            int* o = nodelist + 8 * ni;
            o[0] = at(i, j, k), o[1] = at(i + 1, j, k), o[2] = at(i + 1, j + 1, k), o[3] = at(i, j + 1, k);
            o[4] = at(i, j, k + 1), o[5] = at(i + 1, j, k + 1), o[6] = at(i + 1, j + 1, k + 1), o[7] = at(i, j + 1, k + 1);
            ni++;
          }
    }
    pi += (long)G1 * G1 * G1;
  }
  if (!values) return AMRH_OK;
  int* nodes_dev = nullptr;
  long* z0_dev = nullptr;
  double* out_dev = nullptr;
  hipError_t e = hipMalloc(&nodes_dev, sizeof(int) * order.size());
  if (e == hipSuccess) e = hipMalloc(&z0_dev, sizeof(long) * order.size());
  if (e == hipSuccess) e = hipMalloc(&out_dev, sizeof(double) * std::max<long>(1, nz * c->cfg.nf));
  if (e == hipSuccess)
    e = hipMemcpyAsync(nodes_dev, order.data(), sizeof(int) * order.size(), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(z0_dev, zone0.data(), sizeof(long) * order.size(), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) {
    if (n == 8 && g == 3)
      zone_gather_kernel<8, 3><<<(unsigned)order.size(), 256, 0, c->stream>>>(c->buf[c->cur], c->cfg.nf, nodes_dev,
                                                                              z0_dev, t->covered_dev, nz, out_dev);
    else if (n == 8)
      zone_gather_kernel<8, 2><<<(unsigned)order.size(), 256, 0, c->stream>>>(c->buf[c->cur], c->cfg.nf, nodes_dev,
                                                                              z0_dev, t->covered_dev, nz, out_dev);
    else
      zone_gather_kernel<16, 2><<<(unsigned)order.size(), 256, 0, c->stream>>>(c->buf[c->cur], c->cfg.nf, nodes_dev,
                                                                               z0_dev, t->covered_dev, nz, out_dev);
    e = hipGetLastError();
  }
  if (e == hipSuccess)
    e = hipMemcpyAsync(values, out_dev, sizeof(double) * nz * c->cfg.nf, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  for (void* p : {(void*)nodes_dev, (void*)z0_dev, (void*)out_dev})
    if (p) (void)hipFree(p);
  if (e != hipSuccess) return fail(AMRH_EHIP, "output gather: %s", hipGetErrorString(e));
  return AMRH_OK;
}

int amrh_tree_output_vtk(amrh_ctx* c, const char* path, double time) {
  if (!c || !path) return fail(AMRH_EINVAL, "null argument");
  long npts, nz;
  if (int rc = amrh_tree_output_counts(c, &npts, &nz)) return rc;
  if (nz > 0x7fffffffL / 9 || npts > 0x7fffffffL) return fail(AMRH_EINVAL, "mesh too large for one file");
  const int nf = c->cfg.nf;
  std::vector<double> x(npts), y(npts), z(npts), v((size_t)nz * nf);
  std::vector<int> nl((size_t)nz * 8);
  if (int rc = amrh_tree_output_mesh(c, x.data(), y.data(), z.data(), nl.data(), v.data())) return rc;
  FILE* fp = fopen(path, "wb");
  if (!fp) return fail(AMRH_EINVAL, "cannot open %s for writing", path);
  bool ok = fprintf(fp, "# vtk DataFile Version 3.0\namrhydro time %.17g\nBINARY\nDATASET UNSTRUCTURED_GRID\n", time) > 0;
  ok = ok && fprintf(fp, "POINTS %ld double\n", npts) > 0;
  for (long p = 0; p < npts && ok; p++) {
    put_be(fp, &x[p], 8, &ok);
    put_be(fp, &y[p], 8, &ok);
    put_be(fp, &z[p], 8, &ok);
  }
  ok = ok && fprintf(fp, "\nCELLS %ld %ld\n", nz, nz * 9) > 0;
  const int eight = 8, hex = 12;  // VTK_HEXAHEDRON: the reference's corner order (grid_node.cpp:1333-1340)
  for (long q = 0; q < nz && ok; q++) {
    put_be(fp, &eight, 4, &ok);
    for (int m = 0; m < 8; m++) put_be(fp, &nl[8 * q + m], 4, &ok);
  }
  ok = ok && fprintf(fp, "\nCELL_TYPES %ld\n", nz) > 0;
  for (long q = 0; q < nz && ok; q++) put_be(fp, &hex, 4, &ok);
  ok = ok && fprintf(fp, "\nCELL_DATA %ld\n", nz) > 0;
  // State::field_name (euler/state.cpp:24-42, binary/state_ztwd.cpp:58-89)
  static const char* names[6] = {"d", "sx", "sy", "sz", "et", "tau"};
  static const char* rot_names[13] = {"d_tot", "d_acc", "d_don", "sr", "lz", "sz", "pot",
                                      "et",    "tau",   "T",     "s1", "s2", "s3"};
  for (int l = 0; l < nf && ok; l++) {
    char nm[16];
    if (c->cfg.physics == AMRH_PHYS_EULER && l < 6)
      snprintf(nm, sizeof nm, "%s", names[l]);
    else if (c->cfg.physics == AMRH_PHYS_ROTATING && l < 13)
      snprintf(nm, sizeof nm, "%s", rot_names[l]);
    else
      snprintf(nm, sizeof nm, "f%d", l);
    ok = ok && fprintf(fp, "SCALARS %s double 1\nLOOKUP_TABLE default\n", nm) > 0;
    for (long q = 0; q < nz && ok; q++) put_be(fp, &v[(size_t)l * nz + q], 8, &ok);
    ok = ok && fprintf(fp, "\n") > 0;
  }
  ok = fclose(fp) == 0 && ok;
  if (!ok) return fail(AMRH_EHIP, "writing %s failed", path);
  return AMRH_OK;
}


int amrh_tree_checkpoint_write(amrh_ctx* c, const char* path, double time, double refine_alpha, const int* ages) {
  return amrh_tree_checkpoint_write_ex(c, path, time, refine_alpha, ages, nullptr, 0);
}

int amrh_tree_checkpoint_write_ex(amrh_ctx* c, const char* path, double time, double refine_alpha, const int* ages,
                                  const void* root_prefix, long prefix_bytes) {
  if (!c || !path) return fail(AMRH_EINVAL, "null argument");
  if (prefix_bytes < 0 || (prefix_bytes > 0 && !root_prefix)) return fail(AMRH_EINVAL, "bad root prefix");
  if (!c->tree) return fail(AMRH_ESTATE, "checkpoints hold an octree: call amrh_tree_set");
  const int n = c->cfg.n, g = c->cfg.g, nf = c->cfg.nf, G = n + 2 * g, nb = c->nb;
  const TreeState* t = c->tree;
  std::vector<double> U((size_t)nb * nf * G * G * G);
  if (int rc = amrh_download(c, U.data())) return rc;
  double fo[16] = {0};
  if (int rc = amrh_tree_flow_off(c, fo)) return rc;
  FILE* fp = fopen(path, "wb");
  if (!fp) return fail(AMRH_EINVAL, "cannot open %s for writing", path);
  Writer w{fp};
  if (prefix_bytes > 0 && fwrite(root_prefix, 1, (size_t)prefix_bytes, fp) != (size_t)prefix_bytes) w.ok = false;
  const double h0 = t->h[0];
  std::vector<int> stack = {0};
  while (!stack.empty() && w.ok) {
    const int b = stack.back();
    stack.pop_back();
    w.put(t->h[b]);
    w.put(h0);
    w.put(time);
    for (int d = 0; d < 3; d++) {
      w.put(t->off[b][d]);
      w.put(c->origin[d]);  // every node carries the tree's origin (grid/grid_node.cpp:20-26, :281)
    }
    for (int l = 0; l < nf; l++) w.put(b == 0 ? fo[l] : 0.0);  // only the root keeps a ledger
    const double* Ub = U.data() + (size_t)b * nf * G * G * G;
    for (int i = 0; i < G; i++)
      for (int j = 0; j < G; j++)
        for (int k = 0; k < G; k++)
          for (int l = 0; l < nf; l++) w.put(Ub[(((size_t)l * G + k) * G + j) * G + i]);
    for (int ci = 0; ci < 8; ci++) w.put((bool)(t->child[b][ci] >= 0));
    w.put(refine_alpha);
    w.put(ages ? ages[b] : -1);
    for (int ci = 7; ci >= 0; ci--)
      if (t->child[b][ci] >= 0) stack.push_back(t->child[b][ci]);
  }
  const bool ok = w.ok && fclose(fp) == 0;
  if (!ok) return fail(AMRH_EHIP, "writing %s failed", path);
  return AMRH_OK;
}

int amrh_tree_checkpoint_read(const char* path, const amrh_config* cfg, amrh_ctx** out, int* nblocks,
                              double* time, double* refine_alpha) {
  return amrh_tree_checkpoint_read_ex(path, cfg, out, nblocks, time, refine_alpha, nullptr, 0);
}

int amrh_tree_checkpoint_read_ex(const char* path, const amrh_config* cfg, amrh_ctx** out, int* nblocks,
                                 double* time, double* refine_alpha, void* root_prefix, long prefix_bytes) {
  if (!path || !cfg || !out) return fail(AMRH_EINVAL, "null argument");
  if (prefix_bytes < 0) return fail(AMRH_EINVAL, "bad root prefix");
  *out = nullptr;
  const int n = cfg->n, g = cfg->g, nf = cfg->nf, G = n + 2 * g;
  if (nf <= 0 || nf > 16 || G <= 0) return fail(AMRH_EINVAL, "bad configuration");
  FILE* fp = fopen(path, "rb");
  if (!fp) return fail(AMRH_EINVAL, "cannot open %s", path);
  Reader r{fp};
  if (prefix_bytes > 0) {
    std::vector<unsigned char> pre((size_t)prefix_bytes);
    if (fread(pre.data(), 1, pre.size(), fp) != pre.size()) {
      fclose(fp);
      return fail(AMRH_EINVAL, "%s: shorter than its root prefix", path);
    }
    if (root_prefix) memcpy(root_prefix, pre.data(), pre.size());
  }
  std::vector<int> parent, child;
  std::vector<double> U;
  double fo[16] = {0}, t0 = 0.0, alpha = 1.0, origin[3] = {0.0, 0.0, 0.0};
  // depth first; each frame remembers which child flags are still to be read
  struct Frame {
    int node;
    bool c[8];
    int next;
  };
  std::vector<Frame> stack;
  auto read_node = [&](int par, int ci) -> bool {
    const int b = (int)parent.size();
    parent.push_back(par);
    child.push_back(ci);
    (void)r.get<double>();  // h: implied by the tree
    (void)r.get<double>();  // h0
    const double tm = r.get<double>();
    if (b == 0) t0 = tm;
    for (int d = 0; d < 3; d++) {
      (void)r.get<int>();  // offset: implied by the tree
      const double o = r.get<double>();
      if (b == 0) origin[d] = o;  // the root's; GridNode::set_origin gives every node the same
    }
    for (int l = 0; l < nf; l++) {
      const double v = r.get<double>();
      if (b == 0) fo[l] = v;
    }
    U.resize((size_t)(b + 1) * nf * G * G * G);
    double* Ub = U.data() + (size_t)b * nf * G * G * G;
    for (int i = 0; i < G; i++)
      for (int j = 0; j < G; j++)
        for (int k = 0; k < G; k++)
          for (int l = 0; l < nf; l++) Ub[(((size_t)l * G + k) * G + j) * G + i] = r.get<double>();
    Frame f{b, {}, 0};
    for (int q = 0; q < 8; q++) f.c[q] = r.get<bool>();
    alpha = r.get<double>();
    (void)r.get<int>();  // age
    stack.push_back(f);
    return r.ok;
  };
  bool ok = read_node(-1, -1);
  while (ok && !stack.empty()) {
    Frame& f = stack.back();
    while (f.next < 8 && !f.c[f.next]) f.next++;
    if (f.next == 8) {
      stack.pop_back();
      continue;
    }
    const int par = f.node, ci = f.next++;
    ok = read_node(par, ci);
  }
  fclose(fp);
  if (!ok) return fail(AMRH_EINVAL, "%s: truncated or not a checkpoint of this configuration", path);
  amrh_ctx* c = nullptr;
  const int nb = (int)parent.size();
  if (int rc = amrh_create(cfg, nb, &c)) return rc;
  int rc = amrh_tree_set_origin(c, origin);
  if (!rc) rc = amrh_tree_set(c, parent.data(), child.data());
  if (!rc) rc = amrh_upload(c, U.data());
  if (!rc) {
    hipError_t e = hipMemcpy(c->tree->fo_dev, fo, sizeof(double) * nf, hipMemcpyHostToDevice);
    if (e != hipSuccess) rc = fail(AMRH_EHIP, "ledger upload: %s", hipGetErrorString(e));
  }
  if (rc) {
    amrh_destroy(c);
    return rc;
  }
  if (nblocks) *nblocks = nb;
  if (time) *time = t0;
  if (refine_alpha) *refine_alpha = alpha;
  *out = c;
  return AMRH_OK;
}

}  // extern "C"
