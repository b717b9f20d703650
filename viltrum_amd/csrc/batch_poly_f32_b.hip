/* The batch pipeline for the polynomial family, FP32, dims 4 and 5. */
#include "run_batch.h"

VMX_INSTANTIATE_BATCH(float, 4, vmx::Poly<float>, vmx::FastMath)
VMX_INSTANTIATE_BATCH(float, 5, vmx::Poly<float>, vmx::FastMath)
