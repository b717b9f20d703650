/* The batch pipeline for the polynomial family, FP64, dims 4 and 5. */
#include "run_batch.h"

VMX_INSTANTIATE_BATCH(double, 4, vmx::Poly<double>, vmx::PreciseMath)
VMX_INSTANTIATE_BATCH(double, 5, vmx::Poly<double>, vmx::PreciseMath)
