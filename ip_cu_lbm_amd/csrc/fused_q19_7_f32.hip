// fused_q19_7_f32.hip - fused-path instantiations: D3Q19 flow, D3Q7 scalar, FP32 storage.
#include "fused_kernels.hpp"

using ipl::fusedk::FBufs;
using ipl::fusedk::FCoef;
using ipl::fusedk::FGeom;
using ipl::fusedk::FParams;
using ipl::fusedk::FTables;

IPL_FUSED_INSTANTIATE(float, float, 19, 7)
IPL_FUSED_INSTANTIATE_SCALAR(float, float, 7)
