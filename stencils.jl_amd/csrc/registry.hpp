// Kernel catalogue.  Every build-time specialisation (generated translation
// units) and every runtime-offset kernel registers one entry here at load
// time; sth_map picks the entry that matches the call.
#pragma once

#include <vector>

#include "common.hpp"

namespace sth {

using LaunchFn = void (*)(const MapParams&, hipStream_t);

// Tile configurations of the tiled kernel.
enum TileCfg : int {
  CFG_STATIC2D = 0,  // 64 x 32, shapes fixed at build time
  CFG_STATIC3D = 1,  // 32 x 8 x 8, shapes fixed at build time
  CFG_RT1D = 2,      // 256 x 1 x 1, runtime offsets
  CFG_RT2D = 3,      // 64 x 16 x 1, runtime offsets
  CFG_RT3D = 4,      // 32 x 8 x 4, runtime offsets
  CFG_LIFE = 5,      // SWAR Life kernel (1024 x 4*KY cells per block)
  CFG_STREAM = 6,    // streaming kernel: 64*V x 4*KY cells of one plane per block
  CFG_ZMARCH1 = 7    // z-marching one-step kernel (3-D, reach 1): 64*V x 4*R cells x a z-chunk
};

struct TileDims {
  int tx, ty, tz;
};
inline TileDims tile_dims(int cfg) {
  switch (cfg) {
    case CFG_STATIC2D: return {64, 32, 1};
    case CFG_STATIC3D: return {32, 8, 8};
    case CFG_RT1D: return {256, 1, 1};
    case CFG_RT2D: return {64, 16, 1};
    case CFG_RT3D: return {32, 8, 4};
    default: return {0, 0, 0};
  }
}

struct KernelEntry {
  int dtype;             // sth_dtype of the source
  int op;                // sth_opkind
  int cfg;               // TileCfg
  int L;                 // static shapes: number of offsets (0: runtime)
  int ndim;              // static shapes: stencil dims
  const int8_t* offs;    // static shapes: L x 3 offsets
  LaunchFn fn;
  const char* name;
  int sv = 0;            // streaming / z-marching kernels: cells per lane along dim 1
  int sky = 0;           // streaming kernels: rows per lane; z-marching: rows per wave
};

std::vector<KernelEntry>& registry();

// Kernels that advance an iteration by two steps per launch (temporal
// blocking), matched like KernelEntry.
struct Fused2Entry {
  int dtype;
  int op;
  int L;
  int ndim;
  const int8_t* offs;
  LaunchFn fn;
  int rows;  // 3-D: rows per wave of the z-marching kernel; 2-D: cells per lane
};
std::vector<Fused2Entry>& fused2_registry();

struct Registrar {
  explicit Registrar(const KernelEntry& e) { registry().push_back(e); }
  explicit Registrar(const Fused2Entry& e) { fused2_registry().push_back(e); }
};

}  // namespace sth
