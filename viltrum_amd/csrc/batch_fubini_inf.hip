/* The batch pipeline for the Fubini form over an infinite range (FP32): the Dyson series
   with 1..4 region dims. */
#include "run_batch.h"

using FubiniInf1 = vmx::Fubini<float, vmx::Dyson<float>, VMX_INFINITE_DIM, 1>;
using FubiniInf2 = vmx::Fubini<float, vmx::Dyson<float>, VMX_INFINITE_DIM, 2>;
using FubiniInf3 = vmx::Fubini<float, vmx::Dyson<float>, VMX_INFINITE_DIM, 3>;
using FubiniInf4 = vmx::Fubini<float, vmx::Dyson<float>, VMX_INFINITE_DIM, 4>;
VMX_INSTANTIATE_BATCH(float, 1, FubiniInf1, vmx::FastMath)
VMX_INSTANTIATE_BATCH(float, 2, FubiniInf2, vmx::FastMath)
VMX_INSTANTIATE_BATCH(float, 3, FubiniInf3, vmx::FastMath)
VMX_INSTANTIATE_BATCH(float, 4, FubiniInf4, vmx::FastMath)
