// Compile-time specialised 2D LDS-node kernels (FP32).
#include "kernels2d.h"

Scat2DKernel<float> kym_pick_node_spec_f32(int id) { return pick_node_spec<float>(id); }
Scat2DKernel<float> kym_pick_bwd_node_spec_f32(int id) { return pick_bwd_node_spec<float>(id); }
