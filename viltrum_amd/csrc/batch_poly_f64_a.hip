/* The batch pipeline for the polynomial family, FP64, dims 1..3. */
#include "run_batch.h"

VMX_INSTANTIATE_BATCH(double, 1, vmx::Poly<double>, vmx::PreciseMath)
VMX_INSTANTIATE_BATCH(double, 2, vmx::Poly<double>, vmx::PreciseMath)
VMX_INSTANTIATE_BATCH(double, 3, vmx::Poly<double>, vmx::PreciseMath)
