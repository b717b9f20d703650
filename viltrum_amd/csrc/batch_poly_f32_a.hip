/* The batch pipeline for the polynomial family, FP32, dims 1..3. */
#include "run_batch.h"

VMX_INSTANTIATE_BATCH(float, 1, vmx::Poly<float>, vmx::FastMath)
VMX_INSTANTIATE_BATCH(float, 2, vmx::Poly<float>, vmx::FastMath)
VMX_INSTANTIATE_BATCH(float, 3, vmx::Poly<float>, vmx::FastMath)
