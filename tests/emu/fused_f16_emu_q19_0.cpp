// fused_f16_emu_q19_0.cpp - FP16-storage instantiations of the fused path (D3Q19 flow, scalar lattice 0) for the
// CPU emulation build only (test infrastructure): the library built for the GPU does not carry them.
#include "../../ip_cu_lbm_amd/csrc/fused_kernels.hpp"

using ipl::fusedk::FBufs;
using ipl::fusedk::FCoef;
using ipl::fusedk::FGeom;
using ipl::fusedk::FParams;
using ipl::fusedk::FTables;

IPL_FUSED_INSTANTIATE(__half, float, 19, 0)
IPL_FUSED_INSTANTIATE_POPS(__half, float, 19)
