// The drop-in seam on the reference side: the function a RabbitQCPlus build calls in
// place of care::cpu::correct_cpu (reference include/correct_cpu.hpp:17-21, called at
// src/dispatch_care_correct_cpu.cpp:363). Compiled against the reference's headers and
// linked to libcare_hip.so; see correct_mi355x.cpp.
#ifndef CARE_CORRECT_MI355X_HPP
#define CARE_CORRECT_MI355X_HPP

#include <options.hpp>
#include <cpureadstorage.hpp>
#include <serializedobjectstorage.hpp>

namespace care {
namespace mi355x {

// Corrects every read of `readStorage` on GPU `device` and returns the serialized records
// correct_cpu would return (same bytes; ordered by read id). The GPU builds its own
// minhash tables from the storage, so the CPU minhasher is not needed. Throws
// std::runtime_error with the engine's message when a call fails.
SerializedObjectStorage correct(const ProgramOptions& programOptions, const CpuReadStorage& readStorage,
                                int device = 0);

}  // namespace mi355x
}  // namespace care

#endif
