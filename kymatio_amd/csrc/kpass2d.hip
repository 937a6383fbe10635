// Compile-time specialised 2D big-node pass kernels (FP32).
#include "kernels2d.h"

Scat2DKernel<float> kym_pick_pass_spec_f32(int kind, int id) { return pick_pass_spec<float>(kind, id); }
