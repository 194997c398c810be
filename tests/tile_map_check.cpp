// CPU check of the tiled PhaseField stencil's work decomposition
// (minim_amd/csrc/pf_tile.hpp), run by tests/test_tile_map.py.
//
// For every grid shape and rows-per-thread value below it replays the
// thread -> site mapping of every block and checks that
//  * every x plane the block reads (tile_plane_read) lies inside the grid;
//  * every site of every x plane is owned (stored and summed) exactly once;
//  * every in-plane index a thread loads lies inside the grid;
//  * every LDS cell a thread reads for its y/z neighbours was filled, for the
//    same plane, with exactly the periodic neighbour of the site it computes,
//    and every register neighbour (own rows r +- 1) is that neighbour too.
// Exit status 0 when all shapes pass.
#include <cstdio>
#include <vector>

#include "pf_tile.hpp"

using namespace minim_amd;

static int check_shape(int X, int Y, int Z, int R, int xchunk, int target) {
  TileGeom G;
  if (!tile_geom(X, Y, Z, R, xchunk, target, 8192, G)) {
    std::printf("shape %dx%dx%d R=%d: rejected\n", X, Y, Z, R);
    return 1;
  }
  const long nsite = (long)Y * Z;
  std::vector<int> owned((size_t)X * nsite, 0);
  const int LH = G.TY + 2, LW = kTileZ + 2;
  int bad = 0;
  for (int L = 0; L < G.nblocks && bad < 10; L++) {
    std::vector<long> cell((size_t)LH * LW, -1);  // LDS cell -> in-plane site index
    std::vector<TileThread> th(kTileWaves * kTileZ);
    for (int t = 0; t < kTileWaves * kTileZ; t++) tile_thread(G, L, t, th[t]);
    for (int t = 0; t < kTileWaves * kTileZ; t++) {
      const TileThread& o = th[t];
      if (o.x0 < 0 || o.x1 > X || o.x0 >= o.x1) {
        std::printf("block %d: bad x range [%d,%d)\n", L, o.x0, o.x1);
        return 1;
      }
      // every plane the kernel reads is a plane of the grid, and the planes
      // around each computed plane are its periodic x neighbours
      for (int k = 0; k < tile_planes_read(o.x0, o.x1); k++) {
        const int p = tile_plane_read(o.x0, k, X);
        if (p < 0 || p >= X) {
          if (bad < 5) std::printf("shape %dx%dx%d block %d: plane %d read (k=%d)\n", X, Y, Z, L, p, k);
          bad++;
        }
      }
      for (int i = o.x0; i < o.x1; i++) {
        if (tile_plane(i - 1, X) != (i - 1 + X) % X || tile_plane(i + 1, X) != (i + 1) % X ||
            tile_plane(i + 2, X) != (i + 2) % X)
          bad++;
      }
      for (int r = 0; r < R; r++) {
        if ((long)o.site[r] >= nsite) bad++;
        const int row = o.lrow0 + r;
        if (row < 1 || row > LH - 2 || o.lcol < 1 || o.lcol > LW - 2) bad++;
        long& c = cell[(size_t)row * LW + o.lcol];
        if (c >= 0 && c != (long)o.site[r]) bad++;
        c = o.site[r];
        if (o.own & (1u << r))
          for (int i = o.x0; i < o.x1; i++) owned[(size_t)i * nsite + o.site[r]]++;
      }
      if (o.hsite >= (unsigned)nsite) bad++;
      if (o.hrow >= 0) {
        if (o.hrow >= LH || o.hcol < 0 || o.hcol >= LW) {
          bad++;
          continue;
        }
        long& c = cell[(size_t)o.hrow * LW + o.hcol];
        if (c >= 0 && c != (long)o.hsite) bad++;
        c = o.hsite;
      }
    }
    for (int t = 0; t < kTileWaves * kTileZ; t++) {
      const TileThread& o = th[t];
      for (int r = 0; r < R; r++) {
        if (!(o.own & (1u << r))) continue;
        const long s = o.site[r];
        const long y = s / Z, z = s % Z;
        const long ym = ((y - 1 + Y) % Y) * Z + z, yp = ((y + 1) % Y) * Z + z;
        const long zm = y * Z + (z - 1 + Z) % Z, zp = y * Z + (z + 1) % Z;
        const long got_ym = r > 0 ? (long)o.site[r - 1] : cell[(size_t)(o.lrow0 + r - 1) * LW + o.lcol];
        const long got_yp = r < R - 1 ? (long)o.site[r + 1] : cell[(size_t)(o.lrow0 + r + 1) * LW + o.lcol];
        const long got_zm = cell[(size_t)(o.lrow0 + r) * LW + o.lcol - 1];
        const long got_zp = cell[(size_t)(o.lrow0 + r) * LW + o.lcol + 1];
        if (got_ym != ym || got_yp != yp || got_zm != zm || got_zp != zp) {
          if (bad < 5)
            std::printf("shape %dx%dx%d R=%d block %d thread %d row %d: neighbours %ld %ld %ld %ld want %ld %ld %ld %ld\n",
                        X, Y, Z, R, L, t, r, got_ym, got_yp, got_zm, got_zp, ym, yp, zm, zp);
          bad++;
        }
      }
    }
  }
  for (size_t i = 0; i < owned.size(); i++)
    if (owned[i] != 1) {
      if (bad < 5) std::printf("shape %dx%dx%d R=%d: site %zu owned %d times\n", X, Y, Z, R, i, owned[i]);
      bad++;
    }
  if (bad) std::printf("shape %dx%dx%d R=%d xchunk=%d: %d failures\n", X, Y, Z, R, G.xchunk, bad);
  return bad ? 1 : 0;
}

int main() {
  const int shapes[][3] = {{1, 1, 1},   {1, 1, 7},    {1, 5, 3},    {3, 1, 1},    {2, 2, 2},    {4, 17, 65},
                           {1, 128, 128}, {2, 64, 64}, {2, 1, 1},
                           {5, 16, 64}, {7, 33, 130}, {1, 40, 200}, {9, 3, 129},  {6, 64, 64},  {3, 70, 63},
                           {2, 1, 300}, {11, 9, 1},   {4, 100, 5},  {16, 32, 256}};
  int fails = 0, n = 0;
  for (const auto& s : shapes)
    for (int R = 1; R <= kMaxTileR; R++)
      for (int xc : {0, 1, 3})
        for (int target : {1, 512}) {
          fails += check_shape(s[0], s[1], s[2], R, xc, target);
          n++;
        }
  std::printf("%d cases, %d failed\n", n, fails);
  return fails ? 1 : 0;
}
