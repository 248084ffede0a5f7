// fused_q19_7_f64.hip - fused-path instantiations: D3Q19 flow, D3Q7 scalar, FP64 storage.
#include "fused_kernels.hpp"

using ipl::fusedk::FBufs;
using ipl::fusedk::FCoef;
using ipl::fusedk::FGeom;
using ipl::fusedk::FParams;
using ipl::fusedk::FTables;

IPL_FUSED_INSTANTIATE(double, double, 19, 7)
IPL_FUSED_INSTANTIATE_SCALAR(double, double, 7)
