// pf_tile.hpp -- work decomposition of the tiled PhaseField stencil kernel.
//
// Plain C++ (host and device) so that the mapping can be checked on the CPU:
// every site of the grid must be owned by exactly one thread, and every
// address a thread touches must lie inside the grid.
//
// The canonical grid is X x Y x Z (Z contiguous).  A block of 4 wavefronts
// owns a tile of 64 z-sites x TY = 4R y-rows and marches over a chunk of x
// planes.  Lane l of wavefront w owns column z = z0 + l and the R rows
// y0 + wR .. y0 + wR + R-1 of every plane of the chunk.  Each plane of the
// tile is staged in LDS with a one-site halo ring: LDS row j holds grid row
// y0 - 1 + j and LDS column c holds z0 - 1 + c (rows 0 .. TY+1, columns
// 0 .. 65).  Coordinates are wrapped periodically, so a tile that overhangs
// the grid (Y % TY or Z % 64 != 0, or dimensions smaller than the tile)
// still loads valid sites: its overhanging lanes/rows hold exactly the
// wrapped neighbours the last in-grid sites need, and only "owned" rows are
// stored or summed.
#pragma once

#ifndef MINIM_HD
#if defined(__HIPCC__) || defined(__HIP__)
#define MINIM_HD __host__ __device__
#else
#define MINIM_HD
#endif
#endif

namespace minim_amd {

constexpr int kTileZ = 64;     // sites per tile row: one per lane of a wavefront
constexpr int kTileWaves = 4;  // wavefronts per block, stacked in y
constexpr int kMaxTileR = 4;   // y rows per thread

struct TileGeom {
  int X, Y, Z;
  int R, TY;            // rows per thread, rows per tile (4R)
  int ntz, nty, ntile;  // tiles per plane
  int xchunk, nxc;      // x planes per block, chunks along x
  int nblocks;          // ntile * nxc, x-chunk major
};

struct TileThread {
  int x0, x1;                // planes [x0, x1) marched by this block
  int lrow0, lcol;           // LDS slot of own row 0 (own rows are lrow0 .. lrow0+R-1)
  int hrow, hcol;            // LDS slot this thread fills with a halo site (hrow < 0: none)
  unsigned hsite;            // in-plane index of that halo site (own row 0 when none)
  unsigned site[kMaxTileR];  // in-plane index (y*Z + z) of own rows, wrapped into the grid
  unsigned own;              // bit r: own row r lies inside the grid (stored and summed)
};

MINIM_HD inline int wrap_mod(int v, int n) {
  const int r = v % n;
  return r < 0 ? r + n : r;
}

// x plane of a (possibly out-of-range) plane index v, periodic in X.  The
// kernel reads planes x0-1 .. x1+1 of its chunk, so v lies in [-1, X+1]; with
// X = 1 or 2 that is more than one period away.
MINIM_HD inline int tile_plane(int v, int X) { return wrap_mod(v, X); }

// Every x plane the tiled kernel reads for chunk [x0, x1): the three planes
// x0-1, x0, x0+1 of its prologue and plane i+2 fetched while plane i is
// computed, for i in [x0, x1).
MINIM_HD inline int tile_planes_read(int x0, int x1) { return 3 + (x1 - x0); }
MINIM_HD inline int tile_plane_read(int x0, int k, int X) {
  return k < 3 ? tile_plane(x0 - 1 + k, X) : tile_plane(x0 + 2 + (k - 3), X);
}

// L: logical block index (x-chunk major), t: thread index in the block.
MINIM_HD inline void tile_thread(const TileGeom& G, int L, int t, TileThread& o) {
  const int xc = L / G.ntile, tile = L - xc * G.ntile;
  const int ty_t = tile / G.ntz, tz_t = tile - ty_t * G.ntz;
  const int w = t / kTileZ, lane = t - w * kTileZ;
  const int y0 = ty_t * G.TY, z0 = tz_t * kTileZ;
  const int zl = z0 + lane;
  const int z = wrap_mod(zl, G.Z);
  o.own = 0;
  for (int r = 0; r < kMaxTileR; r++) {
    const int yl = y0 + w * G.R + r;
    o.site[r] = (unsigned)(wrap_mod(yl, G.Y) * G.Z + z);
    if (r < G.R && yl < G.Y && zl < G.Z) o.own |= 1u << r;
  }
  o.x0 = xc * G.xchunk;
  o.x1 = (o.x0 + G.xchunk < G.X) ? o.x0 + G.xchunk : G.X;
  o.lrow0 = w * G.R + 1;
  o.lcol = lane + 1;
  o.hrow = -1;
  o.hcol = 0;
  o.hsite = o.site[0];
  if (t < kTileZ) {  // row above the tile
    o.hrow = 0;
    o.hcol = t + 1;
    o.hsite = (unsigned)(wrap_mod(y0 - 1, G.Y) * G.Z + wrap_mod(z0 + t, G.Z));
  } else if (t < 2 * kTileZ) {  // row below
    o.hrow = G.TY + 1;
    o.hcol = t - kTileZ + 1;
    o.hsite = (unsigned)(wrap_mod(y0 + G.TY, G.Y) * G.Z + wrap_mod(z0 + t - kTileZ, G.Z));
  } else if (t < 2 * kTileZ + G.TY) {  // column before
    const int j = t - 2 * kTileZ;
    o.hrow = j + 1;
    o.hcol = 0;
    o.hsite = (unsigned)(wrap_mod(y0 + j, G.Y) * G.Z + wrap_mod(z0 - 1, G.Z));
  } else if (t < 2 * kTileZ + 2 * G.TY) {  // column after
    const int j = t - 2 * kTileZ - G.TY;
    o.hrow = j + 1;
    o.hcol = kTileZ + 1;
    o.hsite = (unsigned)(wrap_mod(y0 + j, G.Y) * G.Z + wrap_mod(z0 + kTileZ, G.Z));
  }
}

// Tile geometry for R rows per thread.  xchunk > 0 forces the x-march
// length, else the march is sized for about target_blocks blocks.  Returns
// false when the grid cannot be launched (more than max_blocks blocks).
MINIM_HD inline bool tile_geom(int X, int Y, int Z, int R, int xchunk, int target_blocks, int max_blocks,
                               TileGeom& G) {
  G.X = X;
  G.Y = Y;
  G.Z = Z;
  G.R = R;
  G.TY = kTileWaves * R;
  G.ntz = (Z + kTileZ - 1) / kTileZ;
  G.nty = (Y + G.TY - 1) / G.TY;
  G.ntile = G.ntz * G.nty;
  if (xchunk <= 0) {
    long nxc = (target_blocks + G.ntile - 1) / G.ntile;
    if (nxc < 1) nxc = 1;
    if (nxc > X) nxc = X;
    xchunk = (int)((X + nxc - 1) / nxc);
  }
  if (xchunk > X) xchunk = X;
  if (xchunk < 1) xchunk = 1;
  G.xchunk = xchunk;
  G.nxc = (X + xchunk - 1) / xchunk;
  while ((long)G.ntile * G.nxc > max_blocks && G.xchunk < X) {
    G.xchunk *= 2;
    if (G.xchunk > X) G.xchunk = X;
    G.nxc = (X + G.xchunk - 1) / G.xchunk;
  }
  G.nblocks = G.ntile * G.nxc;
  return (long)G.ntile * G.nxc <= max_blocks;
}

}  // namespace minim_amd
