/* The batch pipeline for the polynomial family at 6 dims (the reference's region cap),
   FP32 and FP64. */
#include "run_batch.h"

VMX_INSTANTIATE_BATCH(float, 6, vmx::Poly<float>, vmx::FastMath)
VMX_INSTANTIATE_BATCH(double, 6, vmx::Poly<double>, vmx::PreciseMath)
