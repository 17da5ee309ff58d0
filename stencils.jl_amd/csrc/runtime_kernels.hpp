// Runtime-offset kernels for one storage type: every op the type supports in
// the 1-D, 2-D and 3-D tile configurations.  Included by one generated
// translation unit per type.
#pragma once

#include "registry.hpp"
#include "tile_kernel.hpp"

namespace sth {

template <typename S, int OPK>
void register_runtime_op(const char* name) {
  static constexpr TileDims d1{256, 1, 1}, d2{64, 16, 1}, d3{32, 8, 4};
  registry().push_back(KernelEntry{Elem<S>::code, OPK, CFG_RT1D, 0, 0, nullptr,
                                   &launch_tile<S, OPK, RuntimeShape, d1.tx, d1.ty, d1.tz>, name});
  registry().push_back(KernelEntry{Elem<S>::code, OPK, CFG_RT2D, 0, 0, nullptr,
                                   &launch_tile<S, OPK, RuntimeShape, d2.tx, d2.ty, d2.tz>, name});
  registry().push_back(KernelEntry{Elem<S>::code, OPK, CFG_RT3D, 0, 0, nullptr,
                                   &launch_tile<S, OPK, RuntimeShape, d3.tx, d3.ty, d3.tz>, name});
}

template <typename S>
void register_runtime_type(const char* name) {
  register_runtime_op<S, OP_SUM>(name);
  register_runtime_op<S, OP_MEAN>(name);
  register_runtime_op<S, OP_MIN>(name);
  register_runtime_op<S, OP_MAX>(name);
  if constexpr (!__is_same(S, bool)) register_runtime_op<S, OP_KP>(name);
  if constexpr (Elem<S>::is_float) {
    register_runtime_op<S, OP_LAP>(name);
  } else {
    register_runtime_op<S, OP_LIFE>(name);
  }
}

}  // namespace sth
