/* The batch pipeline for the 5-D lobe family, FP32 and FP64. */
#include "run_batch.h"

VMX_INSTANTIATE_BATCH(float, 5, vmx::Lobes5<float>, vmx::FastMath)
VMX_INSTANTIATE_BATCH(double, 5, vmx::Lobes5<double>, vmx::PreciseMath)
