/* The batch pipeline for the moving-blob family (dims 2 and 3), FP32 and FP64. */
#include "run_batch.h"

VMX_INSTANTIATE_BATCH(float, 2, vmx::Blobs<float>, vmx::FastMath)
VMX_INSTANTIATE_BATCH(float, 3, vmx::Blobs<float>, vmx::FastMath)
VMX_INSTANTIATE_BATCH(double, 2, vmx::Blobs<double>, vmx::PreciseMath)
VMX_INSTANTIATE_BATCH(double, 3, vmx::Blobs<double>, vmx::PreciseMath)
