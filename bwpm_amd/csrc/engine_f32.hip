// FP32 instantiations of the engine (compensated bulk path and the plain diagnostic variant).
#include "engine_impl.hpp"

namespace bwpm {

template class Engine<float, true>;
template class Engine<float, false>;

std::unique_ptr<EngineBase> make_engine_f32(Spec spec, bool compensated) {
  if (compensated) return std::make_unique<Engine<float, true>>(std::move(spec));
  return std::make_unique<Engine<float, false>>(std::move(spec));
}

void run_continuous_lane_f32(LaneFeed& feed, int slots, int chunk, hipStream_t stream, bool compensated) {
  if (compensated)
    Engine<float, true>::continuous(feed, slots, chunk, stream);
  else
    Engine<float, false>::continuous(feed, slots, chunk, stream);
}

}  // namespace bwpm
