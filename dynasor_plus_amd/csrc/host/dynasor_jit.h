// ============================================================================
//  dynasor_jit.h -- the JIT switch of the reference (include/dynasor_jit.h:
//  58-66), kept so callers that consult or set it compile unchanged.
//
//  The device kernels are compiled ahead of time for gfx950, one
//  instantiation per (padded rank, mode count, operand precision), so there
//  is nothing to compile at run time: the switch is recorded and reported,
//  and changes no kernel.  ComputePlan::use_jit is always false.
// ============================================================================
#ifndef DYNASOR_GPU_JIT_H
#define DYNASOR_GPU_JIT_H

namespace dynasor {

// DYN_JIT=1 in the environment, or dyn_jit_force_enable(true), turns it on.
bool dyn_jit_enabled();
void dyn_jit_force_enable(bool on);

}  // namespace dynasor

#endif  // DYNASOR_GPU_JIT_H
