// fused_q19_0_f32.hip - fused-path instantiations: D3Q19 flow, no scalar, FP32 storage.
#include "fused_kernels.hpp"

using ipl::fusedk::FBufs;
using ipl::fusedk::FCoef;
using ipl::fusedk::FGeom;
using ipl::fusedk::FParams;
using ipl::fusedk::FTables;

IPL_FUSED_INSTANTIATE(float, float, 19, 0)
IPL_FUSED_INSTANTIATE_POPS(float, float, 19)
