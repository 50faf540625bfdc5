// FP64 instantiation of the engine (parity mode against the FP64 reference).
#include "engine_impl.hpp"

namespace bwpm {

template class Engine<double, false>;

std::unique_ptr<EngineBase> make_engine_f64(const Spec& spec) { return std::make_unique<Engine<double, false>>(spec); }

}  // namespace bwpm
