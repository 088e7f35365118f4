// Umbrella header of the LLALBM-compatible GPU host layer (the counterpart
// of llalbm/include/llalbm.hpp).  User code includes this header, names the
// HIP policies in its LatticeConfiguration and runs unchanged otherwise.
#ifndef LLALBM_HIP_HPP
#define LLALBM_HIP_HPP

#include "llalbm_hip/aliases.hpp"
#include "llalbm_hip/policies.hpp"
#include "llalbm_hip/io.hpp"
#include "llalbm_hip/flow_analyzer.hpp"
#include "llalbm_hip/lattice.hpp"
#include "llalbm_hip/construction.hpp"

namespace llalbm::core
{
using llalbm::analysis::FlowAnalyzer;
} // namespace llalbm::core

#endif // LLALBM_HIP_HPP
