// fused_q9_9_f32.hip - fused-path instantiations: D2Q9 flow, D2Q9 scalar, FP32 storage.
#include "fused_kernels.hpp"

using ipl::fusedk::FBufs;
using ipl::fusedk::FCoef;
using ipl::fusedk::FGeom;
using ipl::fusedk::FParams;
using ipl::fusedk::FTables;

IPL_FUSED_INSTANTIATE(float, float, 9, 9)
IPL_FUSED_INSTANTIATE_SCALAR(float, float, 9)
