// Checkpoints in the reference's on-disk layout (GridNode::write / read, grid/grid_node.cpp:29-57,
// over Grid::write / read, grid/grid.cpp:104-146), so a run can move between this library and the
// reference.  Depth first from the root, children in index order; per node:
//   double h, double h0, double time, 3 x (int offset, double origin), double FO[nf],
//   the padded GNX^3 state with x slowest and the field fastest ((*this)(i, j, k)[l] for i, j, k, l),
//   8 x bool child-present, double refine_alpha, int age.
// Edge and corner ghost cells are not held by the device; they are written as zeros and ignored
// on reading (the reference never reads them, grid/grid_node.cpp:59-112).
#include <hip/hip_runtime.h>

#include <cstdio>
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

}  // namespace

extern "C" {

int amrh_tree_checkpoint_write(amrh_ctx* c, const char* path, double time, double refine_alpha, const int* ages) {
  if (!c || !path) return fail(AMRH_EINVAL, "null argument");
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
      w.put(0.0);  // origin
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
  if (!path || !cfg || !out) return fail(AMRH_EINVAL, "null argument");
  *out = nullptr;
  const int n = cfg->n, g = cfg->g, nf = cfg->nf, G = n + 2 * g;
  if (nf <= 0 || nf > 16 || G <= 0) return fail(AMRH_EINVAL, "bad configuration");
  FILE* fp = fopen(path, "rb");
  if (!fp) return fail(AMRH_EINVAL, "cannot open %s", path);
  Reader r{fp};
  std::vector<int> parent, child;
  std::vector<double> U;
  double fo[16] = {0}, t0 = 0.0, alpha = 1.0;
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
      (void)r.get<int>();     // offset: implied by the tree
      (void)r.get<double>();  // origin
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
  int rc = amrh_tree_set(c, parent.data(), child.data());
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
