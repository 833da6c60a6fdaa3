// CPU run of the neighbour-list tile kernel's code (simox_amd/csrc/tile.hpp):
// the workgroup-level functions of k_nlist_tile, called for every cell and
// every thread in turn on the host, with the barriers as phase boundaries.
// Every LDS index the kernel forms is checked against its array here, and
// every global index against its buffer (test_tile_cpu.py also builds this
// with the host address sanitizer).
//
//   tile_harness <in> <out>
//   in:  int32 n, nc[3], cap; double box[3], rcut; double x[3n] (cell-sorted);
//        int32 start[ncall + 1]
//   out: int32 err, cnt[n], cnt_fwd[n], col[cap * n] (entry k of atom i at k*n + i)
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../simox_amd/csrc/tile.hpp"

using namespace sx;

// the cell-tile shape under test (staged atoms per LDS chunk, threads)
#ifndef HARNESS_CAP
#define HARNESS_CAP TILE_CAP
#endif
#ifndef HARNESS_NT
#define HARNESS_NT TILE_NT
#endif
constexpr int CAP = HARNESS_CAP, NT = HARNESS_NT;

#define CHECK(c)                                                            \
  do {                                                                      \
    if (!(c)) {                                                             \
      std::fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
      std::exit(3);                                                         \
    }                                                                       \
  } while (0)

template <class T>
static void rd(FILE* f, T* p, size_t k) {
  CHECK(std::fread(p, sizeof(T), k, f) == k);
}

int main(int argc, char** argv) {
  CHECK(argc == 3);
  FILE* f = std::fopen(argv[1], "rb");
  CHECK(f);
  int n, nc[3], cap;
  double box[3], rcut;
  rd(f, &n, 1); rd(f, nc, 3); rd(f, &cap, 1); rd(f, box, 3); rd(f, &rcut, 1);
  const int ncall = nc[0] * nc[1] * nc[2];
  std::vector<double> x(3 * size_t(n));
  std::vector<int> start(ncall + 1);
  rd(f, x.data(), x.size());
  rd(f, start.data(), start.size());
  std::fclose(f);
  CHECK(start[0] == 0 && start[ncall] == n);

  Geom g{};
  g.n = n;
  g.ncall = ncall;
  for (int d = 0; d < 3; ++d) {
    g.nc[d] = nc[d];
    g.box[d] = box[d];
    g.hbox[d] = 0.5 * box[d];
    g.qbox[d] = 0.25 * box[d];
    g.lcell[d] = box[d] / double(nc[d]);
  }
  g.rcut2 = rcut * rcut;
  g.ream2 = g.rcut2;

  std::vector<int> cnt(n, -1), cnt_fwd(n, -1), col(size_t(cap) * n, -1);
  int err = 0;
  TileArgs A{g, x.data(), start.data(), cap, cnt.data(), cnt_fwd.data(), col.data(), &err};
  TileLds<CAP>* L = new TileLds<CAP>;

  for (int b = 0; b < ncall; ++b) {
    int cx, cy, cz, hs, m;
    double h[3];
    tile_home(A, b, cx, cy, cz, h, hs, m);
    CHECK(b >= 0 && b + 1 <= ncall);
    if (m == 0 || m > NPT_MAX) continue;
    for (int t = 0; t < NT; ++t)
      if (t < NVISIT) {
        L->vcell[t] = visit_cell(g, cx, cy, cz, t);
        CHECK(L->vcell[t] >= 0 && L->vcell[t] < ncall);
      }
    for (int t = 0; t < NT; ++t)
      for (int a = t; a < m; a += NT) {
        CHECK(a < NPT_MAX);
        L->kcnt[a] = 0;
        L->kfwd[a] = 0;
      }
    if (!tile_cells_ok(A, *L)) continue;
    for (int v0 = 0; v0 < NVISIT;) {
      const int v1 = tile_chunk_end(A, *L, v0);
      CHECK(v1 > v0 && v1 <= NVISIT);
      tile_layout(A, *L, v0, v1);
      // every LDS slot the chunk uses lies in the position arrays
      CHECK(L->vl[v1] <= CAP);
      for (int v = v0; v < v1; ++v) CHECK(L->vl[v] >= 0 && L->vl[v] <= L->vl[v + 1]);
      // the home cell's visits stage it whole, so slot s0 + a is home atom a
      for (int v = v0; v < v1; ++v)
        if (v == 0 || v == 14) CHECK(L->vl[v + 1] - L->vl[v] == m);
      for (int t = 0; t < NT; ++t) tile_stage(A, *L, v0, v1, t, NT, h);
      const bool wrap = nc[0] < 4 || nc[1] < 4 || nc[2] < 4;
      for (int t = 0; t < NT; ++t)
        for (int a = t; a < m; a += NT)
          wrap ? tile_scan<true, CAP>(A, *L, v0, v1, hs, a, h)
               : tile_scan<false, CAP>(A, *L, v0, v1, hs, a, h);
      v0 = v1;
    }
    for (int t = 0; t < NT; ++t)
      for (int a = t; a < m; a += NT) {
        CHECK(hs + a < n);
        if (tile_finish(A, *L, hs, a) && err == 0) err = SIMOX_ERR_NB_OVERFLOW;
      }
  }
  delete L;
  for (int i = 0; i < n; ++i) {
    CHECK(cnt[i] >= 0 && cnt_fwd[i] >= 0 && cnt_fwd[i] <= cnt[i]);
    for (int k = 0; k < cnt[i]; ++k) {
      const int j = col[size_t(k) * n + i];
      CHECK(j >= 0 && j < n);
    }
  }
  FILE* o = std::fopen(argv[2], "wb");
  CHECK(o);
  std::fwrite(&err, sizeof(int), 1, o);
  std::fwrite(cnt.data(), sizeof(int), n, o);
  std::fwrite(cnt_fwd.data(), sizeof(int), n, o);
  std::fwrite(col.data(), sizeof(int), col.size(), o);
  std::fclose(o);
  return 0;
}
