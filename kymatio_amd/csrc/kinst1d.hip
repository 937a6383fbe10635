// 1D kernels (all dtypes).
#include "kernels1d.h"

Scat1DKernel<float> kym_pick1d_f32(int threads, int kind) { return pick_scat1d_nt<float>(threads, kind); }
Scat1DKernel<double> kym_pick1d_f64(int threads, int kind) { return pick_scat1d_nt<double>(threads, kind); }
Scat1DKernel<float> kym_pick_node1d_spec_f32(int id) { return pick_node1d_spec<float>(id); }
