// FP32 instantiations of the engine (compensated bulk path and the plain diagnostic variant).
#include "engine_impl.hpp"

namespace bwpm {

template class Engine<float, true>;
template class Engine<float, false>;

std::unique_ptr<EngineBase> make_engine_f32(const Spec& spec, bool compensated) {
  if (compensated) return std::make_unique<Engine<float, true>>(spec);
  return std::make_unique<Engine<float, false>>(spec);
}

}  // namespace bwpm
