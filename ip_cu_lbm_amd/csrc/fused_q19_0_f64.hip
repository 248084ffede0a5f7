// fused_q19_0_f64.hip - fused-path instantiations: D3Q19 flow, no scalar, FP64 storage.
#include "fused_kernels.hpp"

using ipl::fusedk::FBufs;
using ipl::fusedk::FCoef;
using ipl::fusedk::FGeom;
using ipl::fusedk::FParams;
using ipl::fusedk::FTables;

IPL_FUSED_INSTANTIATE(double, double, 19, 0)
IPL_FUSED_INSTANTIATE_POPS(double, double, 19)
